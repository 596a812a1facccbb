"""A force field generated from a Yaff parameter file, evaluated on the GPU:
the nonbonded sections of the reference's MIL-53(Al) test input
(yaff/data/test/parameters_mil53.txt with system_mil53.chk: MM3 with 1-2/1-3
exclusions, FIXQ bond-charge-increment charges with Gaussian radii, Ewald
with the generator's alpha = 3.5/rcut, gcut = 1.1 alpha, Switch3 of 4 A on
the MM3 part), against the oracle built from the same parameters: energy,
gradient and virial to 1e-10."""
import os

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason='needs a HIP device')]

HERE = os.path.dirname(os.path.abspath(__file__))
PAR = os.path.join(HERE, 'golden', 'params')
TOL = 1e-10


@pytest.mark.parametrize('tail', [False, True])
def test_mil53_generated_nonbonded_matches_oracle(tail):
    import warnings
    import yaff_amd as ya
    from oracle import oracle as orc
    print('case mil53 generated, tailcorrections=%s' % tail, flush=True)
    system = ya.System.from_file(os.path.join(PAR, 'system_mil53.chk'))
    with warnings.catch_warnings():
        warnings.simplefilter('ignore', RuntimeWarning)
        ff = ya.ForceField.generate(system, os.path.join(PAR, 'parameters_mil53.txt'), skip_other_sections=True,
                                    tailcorrections=tail)
    n = system.natom
    gpos = np.zeros((n, 3))
    vtens = np.zeros((3, 3))
    e = ff.compute(gpos, vtens)

    mm3 = ff.part_pair_mm3.pair_pot
    ei = ff.part_pair_ei.pair_pot
    rc = mm3.rcut
    alpha, gcut = ei.alpha, ff.part_ewald_reci.gcut
    sc_mm3 = ya.Scalings(system, 0.0, 0.0, 1.0)
    sc_ei = ya.Scalings(system, 1.0, 1.0, 1.0)
    O = orc.Oracle()
    cell = O.cell(system.cell.rvecs)
    rows = O.nlist_build(system.pos, cell, rc, orc.rmax_of(cell, rc))
    ref = [O.pair_compute(rows, sc_mm3.stab, orc.Pot(orc.MM3, rc, orc.TR_SWITCH3, mm3.tr.width, p0=mm3.sigmas,
                                                      p1=mm3.epsilons, onlypauli=mm3.onlypaulis), n),
           O.pair_compute(rows, sc_ei.stab, orc.Pot(orc.EI, rc, p0=system.charges, p1=system.radii, alpha=alpha), n),
           O.ewald_reci(system.pos, system.charges, cell, alpha, orc.gmax_of(cell, gcut), gcut),
           O.ewald_corr(system.pos, system.charges, cell, alpha, sc_ei.stab)]
    e_neut = orc.ewald_neutralizing(system.charges, alpha, system.cell.volume, radii=system.radii)
    ref.append((e_neut, np.zeros((n, 3)), -e_neut*np.eye(3)))
    if tail:
        c0, c1 = ff.parts[-1].ecorr, ff.parts[-1].wcorr
        vol = system.cell.volume
        ref.append((2*np.pi*c0/vol, np.zeros((n, 3)), np.eye(3)*2*np.pi*c1/vol))
    e_ref = sum(r[0] for r in ref)
    g_ref = sum(r[1] for r in ref)
    v_ref = sum(r[2] for r in ref)
    assert abs(e - e_ref) <= TOL*abs(e_ref), (e, e_ref)
    assert np.abs(gpos - g_ref).max() <= TOL*np.abs(g_ref).max()
    assert np.abs(vtens - v_ref).max() <= TOL*np.abs(v_ref).max()
