"""Writes the golden fixtures under tests/golden/ (run in a container where
the reference tree and oracle/_ref exist):

    make -C oracle && python tests/golden/make_golden.py

* systems.npz     -- coordinate fixtures of the reference test suite
                     (yaff/test/common.py: water32, water, quartz, graphene8,
                     polyethylene4, caffeine, the 76-atom MIL-53 cell), read
                     as text with the ast module
                     (nothing of the reference is imported or executed), in
                     bohr; quartz is also stored as its 2x2x2 supercell.
* ref_outputs.npz -- outputs of molmod/yaff's own C kernels (oracle/_ref,
                     built by oracle/Makefile from the reference sources) for
                     every case of cases.py: neighbour rows (SHA-256 of the
                     72-byte rows and their count, full rows for small cases),
                     per-potential energy/gpos/vtens, Ewald reciprocal and
                     correction terms.
* ewald_tables.json -- REAL/RECI/CORR energies at five alphas for water32 and
                     quartz as printed in the comments of
                     yaff/pes/test/test_ewald.py:39-45,53-58.  The quartz
                     numbers belong to the 9-atom unit cell (they match it to
                     all printed digits), not to the 2x2x2 supercell the test
                     body now builds.
"""
import ast
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

REF_COMMON = '/root/reference/yaff/test/common.py'
ANGSTROM = 1.889726133921252


def _literal(node):
    """Evaluates np.array(<literal list>) [* int] [* angstrom] parts of the fixture source."""
    if isinstance(node, ast.BinOp) and isinstance(node.op, ast.Add):
        return list(_literal(node.left)) + list(_literal(node.right))
    if isinstance(node, ast.BinOp) and isinstance(node.op, ast.Mult):
        if isinstance(node.right, ast.Name) and node.right.id == 'angstrom':
            return np.asarray(_literal(node.left), float)*ANGSTROM
        left = _literal(node.left)
        right = ast.literal_eval(node.right)
        return list(left)*right
    if isinstance(node, ast.Call) and getattr(node.func, 'attr', None) == 'array':
        return _literal(node.args[0])
    if isinstance(node, ast.List):
        return [_literal(e) for e in node.elts]
    return ast.literal_eval(node)


def extract_systems():
    tree = ast.parse(open(REF_COMMON).read())
    wanted = {'get_system_water32': 'water32', 'get_system_quartz': 'quartz', 'get_system_graphene8': 'graphene8',
              'get_system_polyethylene4': 'polyethylene4', 'get_system_caffeine': 'caffeine',
              'get_system_water': 'water', 'get_system_mil53': 'mil53'}
    out = {}
    for fn in tree.body:
        if not isinstance(fn, ast.FunctionDef) or fn.name not in wanted:
            continue
        name = wanted[fn.name]
        call = [n for n in ast.walk(fn) if isinstance(n, ast.Call) and getattr(n.func, 'id', None) == 'System'][0]
        # keyword values that name a local variable (numbers=numbers) take that variable's assignment
        local = {a.targets[0].id: a.value for a in fn.body
                 if isinstance(a, ast.Assign) and isinstance(a.targets[0], ast.Name)}
        kw = {k.arg: local.get(k.value.id, k.value) if isinstance(k.value, ast.Name) else k.value
              for k in call.keywords}
        out[name + '_numbers'] = np.asarray(_literal(kw['numbers']), np.int64)
        out[name + '_pos'] = np.asarray(_literal(kw['pos']), float)
        if 'rvecs' in kw and not (isinstance(kw['rvecs'], ast.Constant) and kw['rvecs'].value is None):
            out[name + '_rvecs'] = np.asarray(_literal(kw['rvecs']), float)
        if 'charges' in kw:
            out[name + '_charges'] = np.asarray(_literal(kw['charges']), float)
        if 'bonds' in kw:
            if isinstance(kw['bonds'], ast.Call) and isinstance(kw['bonds'].args[0], ast.ListComp):
                # water32: every H bonded to the O that precedes it (i//3*3).
                out[name + '_bonds'] = np.array([[(i//3)*3, i] for i in range(96) if i % 3 != 0], np.int64)
            else:
                out[name + '_bonds'] = np.asarray(_literal(kw['bonds']), np.int64)
    # quartz 2x2x2 supercell (the system of test_ewald.py:53-60)
    from yaff_amd.system import System
    q = System(out['quartz_numbers'], out['quartz_pos'], rvecs=out['quartz_rvecs'], charges=out['quartz_charges'],
               bonds=out['quartz_bonds'])
    q2 = q.supercell(2, 2, 2)
    out['quartz222_numbers'] = q2.numbers
    out['quartz222_pos'] = q2.pos
    out['quartz222_rvecs'] = q2.cell.rvecs
    out['quartz222_charges'] = q2.charges
    out['quartz222_bonds'] = q2.bonds
    return out


def ewald_tables():
    return {
        'water32': {'alpha': [0.05, 0.10, 0.20, 0.50, 1.00],
                    'real': [-7.2121617e-01, -7.1142179e-01, -6.3429485e-01, -7.7158542e-02, -1.9121203e-05],
                    'reci': [2.4919194e-08, 1.5853109e-04, 6.1021373e-03, 9.3260624e-01, 8.2930717e+00],
                    'corr': [-1.4028899e-03, -1.1355778e-02, -9.4426328e-02, -1.5780667e+00, -9.0156717e+00]},
        'quartz': {'alpha': [0.05, 0.10, 0.20, 0.50, 1.00],
                      'real': [-3.5696637e-02, 1.3948043e-01, 1.7482393e-01, 5.6286111e-04, 2.6807119e-12],
                      'reci': [6.8222205e-29, 2.2254505e-08, 2.9254105e-03, 7.7869316e-01, 4.6913539e+00],
                      'corr': [-2.5814534e-01, -4.3332242e-01, -4.7159132e-01, -1.0730980e+00, -4.9851959e+00]},
    }


def rows_digest(rows):
    return hashlib.sha256(np.ascontiguousarray(rows).tobytes()).hexdigest()


def reference_outputs(cases):
    from oracle.oracle import Oracle, rmax_of, gmax_of
    import cases as hg
    R = Oracle(reference=True)
    out = {}
    for name, c in cases.items():
        cell = R.cell(c['rvecs'])
        rmax = rmax_of(cell, c['rc']) if cell.nvec else np.zeros(0, np.int64)
        rows = R.nlist_build(c['pos'], cell, c['rc'], rmax)
        out[name + '/nrows'] = np.array(len(rows))
        out[name + '/rows_sha256'] = np.array(rows_digest(rows))
        if len(rows) <= 20000:
            out[name + '/rows'] = rows
        n = len(c['pos'])
        for ip, p in enumerate(c['pots']):
            e, g, v = R.pair_compute(rows, c['stab'], hg.oracle_pot(p), n)
            out['%s/pot%d/e' % (name, ip)] = np.array(e)
            out['%s/pot%d/gpos' % (name, ip)] = g
            out['%s/pot%d/vtens' % (name, ip)] = v
            e2, _, _ = R.pair_compute(rows, c['stab'], hg.oracle_pot(p), n, False, False)
            out['%s/pot%d/e_nograd' % (name, ip)] = np.array(e2)
        if c['ewald'] is not None:
            ew = c['ewald']
            gmax = gmax_of(cell, ew['gcut'])
            e, g, v = R.ewald_reci(c['pos'], c['charges'], cell, ew['alpha'], gmax, ew['gcut'],
                                   ew.get('dielectric', 1.0))
            out[name + '/reci/e'] = np.array(e)
            out[name + '/reci/gpos'] = g
            out[name + '/reci/vtens'] = v
            out[name + '/reci/gmax'] = gmax
            e, g, v = R.ewald_corr(c['pos'], c['charges'], cell, ew['alpha'], c['stab'], ew.get('dielectric', 1.0))
            out[name + '/corr/e'] = np.array(e)
            out[name + '/corr/gpos'] = g
            out[name + '/corr/vtens'] = v
    # Cell golden vectors of test_cell.py:38-56 through the reference C.
    cell = R.cell(cases['water32']['rvecs'])
    out['cell/mic'] = R.mic(np.array([10.0, 0.0, 5.0])*ANGSTROM, cell)
    out['cell/add_vec'] = R.add_vec(out['cell/mic'], cell, np.array([1, 2, 3]))
    # Switch3 / Hammer values at the points of test_truncation.py.
    pts = []
    for (kind, par) in ((1, 1.0), (1, 0.5), (2, 1.0), (2, 0.5)):
        for d in (0.0, 8.0, 9.0, 9.5, 9.99, 10.0, 10.5):
            h, g = R.trunc(kind, par, d, 10.0)
            pts.append((kind, par, d, h, g))
    out['trunc'] = np.array(pts)
    return out


def main():
    sysd = extract_systems()
    np.savez_compressed(os.path.join(HERE, 'systems.npz'), **sysd)
    with open(os.path.join(HERE, 'ewald_tables.json'), 'w') as f:
        json.dump(ewald_tables(), f, indent=1)
    import cases as cs
    out = reference_outputs(cs.build_cases())
    np.savez_compressed(os.path.join(HERE, 'ref_outputs.npz'), **out)
    print('wrote', len(out), 'reference arrays')


if __name__ == '__main__':
    main()
