"""Writes tailcorr_ref.json: the (C0, C1) tail-correction sums of the cases in
tailcorr_cases.py from molmod/yaff's own C (pair_pot_tailcorr_cut and
pair_pot_tailcorr_switch3, yaff/pes/pair_pot.c:178-254, with tailcorr.c)
compiled into oracle/_ref by oracle/Makefile.

    make -C oracle && python tests/golden/make_tailcorr_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

from oracle import oracle as orc  # noqa: E402
import tailcorr_cases  # noqa: E402


def ref_pot(c):
    tr = orc.TR_SWITCH3 if c['width'] is not None else orc.TR_NONE
    par = c['width'] or 0.0
    if c['kind'] == 'lj':
        return orc.Pot(orc.LJ, c['rcut'], tr, par, p0=c['sigmas'], p1=c['epsilons'])
    if c['kind'] == 'mm3':
        return orc.Pot(orc.MM3, c['rcut'], tr, par, p0=c['sigmas'], p1=c['epsilons'], onlypauli=c['onlypaulis'])
    return orc.Pot(orc.EXPREP, c['rcut'], tr, par, ffatype_ids=c['ffatype_ids'], amp_cross=c['amp_cross'],
                   b_cross=c['b_cross'])


def main():
    O = orc.Oracle(reference=True)
    out = [list(O.tailcorr(ref_pot(c), c['natom'], c['nlow'], c['nhigh'])) for c in tailcorr_cases.build_cases()]
    with open(os.path.join(HERE, 'tailcorr_ref.json'), 'w') as f:
        json.dump({'source': 'oracle/_ref/libyaffref.so pair_pot_tailcorr_cut/_switch3', 'c0_c1': out}, f, indent=0)
    print('wrote %d cases' % len(out))


if __name__ == '__main__':
    main()
