"""qcf -- MI355X-native Pauli-frame sampling and bit-packed Pauli products.

Host-side mirror of QuantumClifford.jl's interface for its frame-sampling hot
path (`pftrajectories`, `PauliFrame`, `measurements`) and `mul_left!`, over
the C ABI of libqcf.so (include/qcf.h).  Every computation runs in the HIP
library; nothing here computes results on the CPU.
"""
from ._lib import (ArgumentError, BoundsError, DimensionMismatch, HIPError, MethodError,
                   QCFError, lib)
from .ops import *  # noqa: F401,F403  (reference operation names)
from .ops import Depolarize2, PauliError, affectedbits, affectedqubits
from .pauli import (P, PauliOperator, S, Tableau, canonicalize_, mul_left_,
                    mul_left_destabilizer_, mul_left_pauli_row_, mul_left_row_, mul_left_sweep_,
                    mul_left_tableau_pauli_, mul_left_tableau_row_, mul_left_vec_, mul_right_,
                    mul_right_pauli_row_, mul_right_row_, mul_right_tableau_pauli_,
                    mul_right_tableau_row_)
from .frames import (PauliFrame, Program, alloc_masks, apply_, applynoise_,
                     frame_from_reference, frame_to_reference, initZ_, lower_circuit,
                     measurements, pfmeasurements, pftrajectories, pftrajectories_state, seed_,
                     view)
from . import batch, circuits, decoders, jltext, randoms

__all__ = [n for n in dir() if not n.startswith("_")]
