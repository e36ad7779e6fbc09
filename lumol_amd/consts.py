"""Physical constants in lumol's internal units (lumol-core/src/consts.rs)."""
K_BOLTZMANN = 8.31446284161522e-7
BOHR_RADIUS = 0.52917720859
AVOGADRO_NUMBER = 6.02214179e23
FOUR_PI_EPSILON_0 = 7.197589831304046
