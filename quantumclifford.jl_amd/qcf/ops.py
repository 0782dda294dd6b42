"""Circuit operations with the reference's names and constructor semantics.

Each class mirrors a QuantumClifford.jl struct (file:line of the reference
given per group) and lowers to one or more `qcf_op` records of include/qcf.h.
Qubits and bits are 1-based as in the reference.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Iterable, Sequence

from ._lib import ArgumentError

# enum qcf_op_kind (include/qcf.h)
K = dict(
    NOP=0, ID1=1, HADAMARD=2, HADAMARD_XY=3, HADAMARD_YZ=4, PHASE=5, INV_PHASE=6, X=7, Y=8,
    Z=9, SQRTX=10, INV_SQRTX=11, SQRTY=12, INV_SQRTY=13, CXYZ=14, CZYX=15,
    SINGLE_QUBIT_OPERATOR=20,
    SWAP=32, SWAPCX=33, INV_SWAPCX=34, ISWAP=35, INV_ISWAP=36, CZSWAP=37, CXSWAP=38, CNOT=39,
    CPHASE=40, ZCX=41, ZCY=42, ZCZ=43, XCX=44, XCY=45, XCZ=46, YCX=47, YCY=48, YCZ=49,
    ZCRY=50, INV_ZCRY=51, SQRTZZ=52, INV_SQRTZZ=53, SQRTXX=54, INV_SQRTXX=55, SQRTYY=56,
    INV_SQRTYY=57,
    MZ=64, MRZ=65, MX=66, MRX=67, PAULI_MEASUREMENT=68,
    PAULI_ERROR=80, PAULI_NOISE=81, DEPOLARIZE2=82,
    CLASSICAL_XOR=96,
)


class Operation:
    """Base class; subclasses append qcf_op tuples through `lower`."""

    def lower(self, emit, args: list, nqubits: int) -> None:  # pragma: no cover
        raise NotImplementedError

    def affectedqubits(self) -> tuple:
        return ()

    def affectedbits(self) -> tuple:
        return ()


def _nozero(*qs):
    for q in qs:
        if q <= 0:
            raise ArgumentError("Qubit indices have to be larger than zero.")


# ---- symbolic single-qubit gates (src/symbolic_cliffords.jl:105-130) -----
class AbstractSingleQubitOperator(Operation):
    kind = None

    def __init__(self, q: int):
        _nozero(q)
        self.q = int(q)

    def lower(self, emit, args, nqubits):
        emit(K[self.kind], self.q)

    def affectedqubits(self):
        return (self.q,)

    def __repr__(self):
        return f"{type(self).__name__}({self.q})"


def _single(name, kind):
    return type(name, (AbstractSingleQubitOperator,), {"kind": kind})


sId1 = _single("sId1", "ID1")
sHadamard = _single("sHadamard", "HADAMARD")
sHadamardXY = _single("sHadamardXY", "HADAMARD_XY")
sHadamardYZ = _single("sHadamardYZ", "HADAMARD_YZ")
sPhase = _single("sPhase", "PHASE")
sInvPhase = _single("sInvPhase", "INV_PHASE")
sX = _single("sX", "X")
sY = _single("sY", "Y")
sZ = _single("sZ", "Z")
sSQRTX = _single("sSQRTX", "SQRTX")
sInvSQRTX = _single("sInvSQRTX", "INV_SQRTX")
sSQRTY = _single("sSQRTY", "SQRTY")
sInvSQRTY = _single("sInvSQRTY", "INV_SQRTY")
sCXYZ = _single("sCXYZ", "CXYZ")
sCZYX = _single("sCZYX", "CZYX")


class SingleQubitOperator(AbstractSingleQubitOperator):
    """SingleQubitOperator(q, xx, xz, zx, zz, px, pz) (src/symbolic_cliffords.jl:168-206)."""

    kind = "SINGLE_QUBIT_OPERATOR"

    def __init__(self, q, xx, xz, zx, zz, px=False, pz=False):
        super().__init__(q)
        self.m = (bool(xx), bool(xz), bool(zx), bool(zz))
        self.px, self.pz = bool(px), bool(pz)

    def lower(self, emit, args, nqubits):
        off = len(args)
        args.extend(int(v) for v in self.m)
        emit(K[self.kind], self.q, arg_off=off, arg_len=4)


# ---- symbolic two-qubit gates (src/symbolic_cliffords.jl:344-395) --------
class AbstractTwoQubitOperator(Operation):
    kind = None

    def __init__(self, q1: int, q2: int):
        _nozero(q1, q2)
        if q1 == q2:
            raise ArgumentError(
                "Failed to create a two qubit gate because the two qubits it acts on have "
                "the same index. The qubit indices have to be different.")
        self.q1, self.q2 = int(q1), int(q2)

    def lower(self, emit, args, nqubits):
        emit(K[self.kind], self.q1, self.q2)

    def affectedqubits(self):
        return (self.q1, self.q2)

    def __repr__(self):
        return f"{type(self).__name__}({self.q1},{self.q2})"


def _two(name, kind):
    return type(name, (AbstractTwoQubitOperator,), {"kind": kind})


sSWAP = _two("sSWAP", "SWAP")
sSWAPCX = _two("sSWAPCX", "SWAPCX")
sInvSWAPCX = _two("sInvSWAPCX", "INV_SWAPCX")
sISWAP = _two("sISWAP", "ISWAP")
sInvISWAP = _two("sInvISWAP", "INV_ISWAP")
sCZSWAP = _two("sCZSWAP", "CZSWAP")
sCXSWAP = _two("sCXSWAP", "CXSWAP")
sCNOT = _two("sCNOT", "CNOT")
sCPHASE = _two("sCPHASE", "CPHASE")
sZCX = _two("sZCX", "ZCX")
sZCY = _two("sZCY", "ZCY")
sZCZ = _two("sZCZ", "ZCZ")
sXCX = _two("sXCX", "XCX")
sXCY = _two("sXCY", "XCY")
sXCZ = _two("sXCZ", "XCZ")
sYCX = _two("sYCX", "YCX")
sYCY = _two("sYCY", "YCY")
sYCZ = _two("sYCZ", "YCZ")
sZCrY = _two("sZCrY", "ZCRY")
sInvZCrY = _two("sInvZCrY", "INV_ZCRY")
sSQRTZZ = _two("sSQRTZZ", "SQRTZZ")
sInvSQRTZZ = _two("sInvSQRTZZ", "INV_SQRTZZ")
sSQRTXX = _two("sSQRTXX", "SQRTXX")
sInvSQRTXX = _two("sInvSQRTXX", "INV_SQRTXX")
sSQRTYY = _two("sSQRTYY", "SQRTYY")
sInvSQRTYY = _two("sInvSQRTYY", "INV_SQRTYY")

SINGLE_QUBIT_GATES = [sId1, sHadamard, sHadamardXY, sHadamardYZ, sPhase, sInvPhase, sX, sY, sZ,
                      sSQRTX, sInvSQRTX, sSQRTY, sInvSQRTY, sCXYZ, sCZYX]
TWO_QUBIT_GATES = [sSWAP, sSWAPCX, sInvSWAPCX, sISWAP, sInvISWAP, sCZSWAP, sCXSWAP, sCNOT,
                   sCPHASE, sZCX, sZCY, sZCZ, sXCX, sXCY, sXCZ, sYCX, sYCY, sYCZ, sZCrY,
                   sInvZCrY, sSQRTZZ, sInvSQRTZZ, sSQRTXX, sInvSQRTXX, sSQRTYY, sInvSQRTYY]


# ---- measurements (src/symbolic_cliffords.jl:515-540, 626-655) -----------
class _Measure(Operation):
    kind = None

    def __init__(self, qubit: int, bit: int | None = 0):
        _nozero(qubit)
        self.qubit = int(qubit)
        self.bit = 0 if bit is None else int(bit)

    def lower(self, emit, args, nqubits):
        emit(K[self.kind], self.qubit, bit=self.bit)

    def affectedqubits(self):
        return (self.qubit,)

    def affectedbits(self):
        return () if self.bit == 0 else (self.bit,)

    def __repr__(self):
        return f"{type(self).__name__}({self.qubit},{self.bit})"


class sMZ(_Measure):
    kind = "MZ"


class sMRZ(_Measure):
    kind = "MRZ"


class sMX(_Measure):
    kind = "MX"


class sMRX(_Measure):
    kind = "MRX"


class sMY(_Measure):
    """No PauliFrame method in the reference: lowering raises MethodError."""
    kind = "MY"

    def lower(self, emit, args, nqubits):
        from ._lib import MethodError
        raise MethodError("no method matching apply!(::PauliFrame, ::sMY)")


class sMRY(sMY):
    kind = "MRY"


# ---- noise (src/noise.jl:26-199) ------------------------------------------
@dataclass(frozen=True)
class UnbiasedUncorrelatedNoise:
    p: float


@dataclass(frozen=True)
class PauliNoise:
    px: float
    py: float
    pz: float


def PauliNoise_(p):  # PauliNoise(p) constructs UnbiasedUncorrelatedNoise(p) (noise.jl:69-72)
    return UnbiasedUncorrelatedNoise(float(p))


@dataclass(frozen=True)
class DepolarizationNoise:
    """DepolarizationNoise(p): with probability p a uniformly random n-qubit Pauli,
    identity included (src/noise.jl:31-53).  On frames it is lowered jointly: one
    qubit -> X/Y/Z each p/4; two qubits -> each of the 15 non-identity Paulis p/16."""
    p: float

    @staticmethod
    def two_qubit(e: float) -> "DepolarizationNoise":
        return DepolarizationNoise(e * 16 / 15)


class NoiseOp(Operation):
    def __init__(self, noise, indices: Iterable[int]):
        self.noise = noise
        self.indices = tuple(int(i) for i in indices)
        _nozero(*self.indices)

    def affectedqubits(self):
        return self.indices

    def lower(self, emit, args, nqubits):
        _lower_noise(self.noise, self.indices, emit)

    def __repr__(self):
        return f"NoiseOp({self.noise}, {self.indices})"


def _lower_noise(noise, indices, emit):
    if isinstance(noise, UnbiasedUncorrelatedNoise):
        for q in indices:
            emit(K["PAULI_ERROR"], q, p=(noise.p, 0.0, 0.0))
    elif isinstance(noise, PauliNoise):
        for q in indices:
            emit(K["PAULI_NOISE"], q, p=(noise.px, noise.py, noise.pz))
    elif isinstance(noise, DepolarizationNoise):
        if len(indices) == 1:
            e = 0.75 * noise.p
            emit(K["PAULI_ERROR"], indices[0], p=(e, 0.0, 0.0))
        elif len(indices) == 2:
            emit(K["DEPOLARIZE2"], indices[0], indices[1], p=(noise.p * 15 / 16, 0.0, 0.0))
        else:
            from ._lib import MethodError
            raise MethodError("DepolarizationNoise on frames supports one or two qubits")
    else:
        from ._lib import MethodError
        raise MethodError(f"no applynoise! method for {type(noise).__name__} on PauliFrame")


def PauliError(qubits, p, py=None, pz=None):
    """PauliError(q, p) / PauliError(qubits, p) / PauliError(q, px, py, pz)
    (src/noise.jl:146-171)."""
    qs = (qubits,) if isinstance(qubits, int) else tuple(qubits)
    if py is None:
        return NoiseOp(UnbiasedUncorrelatedNoise(float(p)), qs)
    return NoiseOp(PauliNoise(float(p), float(py), float(pz)), qs)


class Depolarize2(NoiseOp):
    """Two-qubit depolarizing channel on frames: with probability e one of the 15
    non-identity two-qubit Paulis, uniformly.  The reference has no PauliFrame
    method for a joint two-qubit channel; this is the distribution of
    DepolarizationNoise(two_qubit_error_probability=e) (src/noise.jl:49-53, 99-129)."""

    def __init__(self, q1: int, q2: int, e: float):
        if q1 == q2:
            raise ArgumentError("Depolarize2 needs two different qubits")
        super().__init__(DepolarizationNoise.two_qubit(float(e)), (q1, q2))
        self.e = float(e)

    def lower(self, emit, args, nqubits):
        emit(K["DEPOLARIZE2"], self.indices[0], self.indices[1], p=(self.e, 0.0, 0.0))


class NoiseOpAll(Operation):
    """NoiseOpAll(noise): noise on qubits 1..nqubits(frame) (src/noise.jl:174-199)."""

    def __init__(self, noise):
        self.noise = noise

    def lower(self, emit, args, nqubits):
        _lower_noise(self.noise, tuple(range(1, nqubits + 1)), emit)


class NoisyGate(Operation):
    """NoisyGate(gate, noise): gate, then noise on its qubits (src/noise.jl:183-195)."""

    def __init__(self, gate, noise):
        self.gate, self.noise = gate, noise

    def affectedqubits(self):
        return self.gate.affectedqubits()

    def affectedbits(self):
        return self.gate.affectedbits()

    def lower(self, emit, args, nqubits):
        self.gate.lower(emit, args, nqubits)
        _lower_noise(self.noise, tuple(self.gate.affectedqubits()), emit)


# ---- classical and Pauli measurements (src/misc_ops.jl) -------------------
class ClassicalXOR(Operation):
    """ClassicalXOR(bits, store) (src/misc_ops.jl:198-205)."""

    def __init__(self, bits: Sequence[int], store: int):
        self.bits = tuple(int(b) for b in bits)
        self.store = int(store)

    def affectedbits(self):
        return (*self.bits, self.store)

    def lower(self, emit, args, nqubits):
        off = len(args)
        args.extend(self.bits)
        emit(K["CLASSICAL_XOR"], 0, bit=self.store, arg_off=off, arg_len=len(self.bits))

    def __repr__(self):
        return f"ClassicalXOR({self.bits},{self.store})"


class PauliMeasurement(Operation):
    """PauliMeasurement(pauli, bit) (src/misc_ops.jl:9-17), applied to frames through
    the ancilla nqubits(pauli)+1 (src/pauli_frames.jl:118-134)."""

    def __init__(self, pauli, bit: int | None = 0):
        self.pauli = pauli
        self.bit = 0 if bit is None else int(bit)

    def affectedqubits(self):
        return tuple(range(1, self.pauli.nqubits + 1))

    def affectedbits(self):
        return () if self.bit == 0 else (self.bit,)

    def lower(self, emit, args, nqubits):
        off = len(args)
        xs, zs = self.pauli.xbits(), self.pauli.zbits()
        args.extend(int(x) | (int(z) << 1) for x, z in zip(xs, zs))
        args.append(int(self.pauli.phase_value()))
        emit(K["PAULI_MEASUREMENT"], 0, bit=self.bit, arg_off=off,
             arg_len=self.pauli.nqubits + 1)


def affectedqubits(op) -> tuple:
    return op.affectedqubits()


def affectedbits(op) -> tuple:
    return op.affectedbits()
