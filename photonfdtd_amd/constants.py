"""SI constants used by the scene description and the stepper.

Values are the ones photonfdtd fixes in ``src/photonfdtd/constants.py:3-8``;
they must agree to the last bit because every update coefficient
(``dt/(eps*EPS_0)``, ``dt/MU_0``, CPML ``sigma``) is derived from them.
"""
import math

C_0 = 299_792_458.0                 # speed of light in vacuum, m/s
EPS_0 = 8.854_187_8128e-12          # vacuum permittivity, F/m
MU_0 = 1.256_637_062_12e-6          # vacuum permeability, H/m
ETA_0 = math.sqrt(MU_0 / EPS_0)     # vacuum wave impedance, ohm
Q_E = 1.602_176_634e-19             # elementary charge, C
