"""Single faults and gauge draws of a compiled program, pushed through the
oracle one per shot via replay masks, and the exact record marginals they
determine.

Replay planes are laid out as libqcf documents them: the initZ! z words of
qubits 1..nq, then one plane per sMRZ/sMRX z-flip (and per PauliMeasurement,
whose ancilla is measured and reset), two (x, z) per
single-qubit noise site and four (x1, z1, x2, z2) per two-qubit depolarizing
site, in op order.  The Pauli choices of a site are mutually exclusive
(src/pauli_frames.jl:136-174; the two-qubit channel follows
DepolarizationNoise, src/noise.jl:99-129: uniform over the 15 non-identity
Paulis, first qubit in the least significant base-4 digit, 1=X 2=Z 3=Y);
distinct sites are independent.
"""
import numpy as np

import oracle
import qcf
from qcf.ops import K

FAULTS_1Q = {1: (1, 0), 2: (0, 1), 3: (1, 1)}  # X, Z, Y -> (x flip, z flip)


def program_faults(prog):
    """([(site, planes, rate)], gauge_planes): every Pauli choice of every noise
    site with the replay planes it flips, and the planes of gauge draws."""
    faults, gauge = [], list(range(prog.nq))
    plane, site = prog.nq, 0
    for op in prog.ops:
        k = int(op["kind"])
        if k in (K["MRZ"], K["MRX"], K["PAULI_MEASUREMENT"]):  # the last resets its ancilla
            gauge.append(plane)
            plane += 1
        elif k == K["PAULI_ERROR"]:
            p = float(op["p"][0])
            for pa in (1, 2, 3):
                faults.append((site, [plane + j for j in range(2) if FAULTS_1Q[pa][j]], p / 3))
            plane, site = plane + 2, site + 1
        elif k == K["PAULI_NOISE"]:
            px, py, pz = (float(x) for x in op["p"])
            for pa, q in ((1, px), (2, pz), (3, py)):
                if q > 0:
                    faults.append((site, [plane + j for j in range(2) if FAULTS_1Q[pa][j]], q))
            plane, site = plane + 2, site + 1
        elif k == K["DEPOLARIZE2"]:
            e = float(op["p"][0])
            for pa in range(1, 16):
                d1, d2 = pa & 3, pa >> 2
                pl = [plane + j for j in range(2) if d1 and FAULTS_1Q[d1][j]]
                pl += [plane + 2 + j for j in range(2) if d2 and FAULTS_1Q[d2][j]]
                faults.append((site, pl, e / 15))
            plane, site = plane + 4, site + 1
    assert plane == prog.nreplay
    return faults, gauge


def inject(prog, plane_sets):
    """Oracle record flips (nbits, len(plane_sets)): shot s replays exactly the
    planes plane_sets[s] (every other draw zero) on a fresh frame."""
    shots = len(plane_sets)
    bits = np.zeros((prog.nreplay, shots), np.uint8)
    for s, pl in enumerate(plane_sets):
        bits[pl, s] = 1
    _, meas = oracle.pf_run(prog.ops, prog.args, prog.nq, prog.nbits, shots,
                            masks=oracle.pack_shots(bits), nthreads=4)
    return meas[: prog.nbits]


def exact_record_marginals(prog):
    """P(record bit b = 1) for every bit of a fresh native run: 1/2 where any
    gauge draw reaches the bit (it is then XORed with an independent uniform
    bit), else the XOR of independent sites, (1 - prod_s (1 - 2 f_s)) / 2 with
    f_s the total rate of site s's Pauli choices that flip the bit."""
    faults, gauge = program_faults(prog)
    fm = inject(prog, [pl for _, pl, _ in faults]).astype(bool)
    gm = inject(prog, [[g] for g in gauge]).astype(bool)
    nsites = 1 + max((s for s, _, _ in faults), default=-1)
    f = np.zeros((nsites, prog.nbits))
    for j, (s, _, p) in enumerate(faults):
        f[s, fm[:, j]] += p
    prob = 0.5 * (1.0 - np.prod(1.0 - 2.0 * f, axis=0))
    prob[gm.any(axis=1)] = 0.5
    return prob
