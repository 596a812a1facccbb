"""Atomic units used throughout (bohr, hartree, electron mass, atomic time).

The angstrom is the value Yaff's generator defaults are written in
(yaff/pes/generator.py:90: 10 A = 18.89726133921252 bohr).  The other
constants follow molmod's CODATA-2006 based definitions; only MD
trajectories depend on them, never force parity.
"""
angstrom = 1.889726133921252
amu = 1.66053886e-27/9.10938215e-31
femtosecond = 1e-15/2.418884326505e-17
boltzmann = 1.3806504e-23/4.35974394e-18
kelvin = 1.0
kjmol = 1000.0/(4.35974394e-18*6.02214179e23)
