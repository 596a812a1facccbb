"""Tail-correction parity cases (inputs only, seeded): LJ, MM3 (with and
without onlypauli) and ExpRep, each with a plain cutoff and a Switch3, over
all atoms and with nlow/nhigh exclusions.  Their reference outputs -- the
(C0, C1) sums of molmod/yaff's pair_pot_tailcorr_cut / _switch3 compiled into
oracle/_ref -- are written to tailcorr_ref.json by make_tailcorr_golden.py."""
import numpy as np

ANGSTROM = 1.889726133921252
KCALMOL = 0.0015936015865590541  # Hartree per kcal/mol


def build_cases():
    rng = np.random.default_rng(11)
    natom = 37
    cases = []
    sig = 2.1*ANGSTROM + rng.normal(0.0, 0.05*ANGSTROM, natom)
    eps = 0.15*KCALMOL + rng.normal(0.0, 0.005*KCALMOL, natom)
    ops = (rng.uniform(0, 1, natom) < 0.3).astype(np.int32)
    types = rng.integers(0, 3, natom)
    # one type pair without repulsion (amp 0); every b nonzero: for b == 0 the
    # reference's closed form divides by b (NaN) where its pair function
    # returns zero, and this build follows the pair function
    amp = np.array([[2.0, 1.5, 0.7], [1.5, 3.0, 0.0], [0.7, 0.0, 1.1]])
    bb = np.array([[3.1, 2.9, 2.5], [2.9, 3.3, 2.0], [2.5, 2.0, 2.2]])/ANGSTROM
    for rc, w in ((12.0*ANGSTROM, None), (12.0*ANGSTROM, 3.0*ANGSTROM), (9.0*ANGSTROM, 4.0*ANGSTROM)):
        for nlow, nhigh in ((0, natom), (5, natom), (0, 30), (7, 22)):
            base = dict(rcut=rc, width=w, natom=natom, nlow=nlow, nhigh=nhigh)
            cases.append(dict(base, kind='lj', sigmas=sig, epsilons=eps))
            cases.append(dict(base, kind='mm3', sigmas=sig, epsilons=eps, onlypaulis=ops))
            cases.append(dict(base, kind='mm3', sigmas=sig, epsilons=eps, onlypaulis=np.zeros(natom, np.int32)))
            cases.append(dict(base, kind='exprep', ffatype_ids=types, amp_cross=amp, b_cross=bb))
    return cases
