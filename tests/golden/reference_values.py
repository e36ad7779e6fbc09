"""Known answers held by the reference's own tests, transcribed as data.

Every value cites the reference file:line that asserts it (paths relative to
the reference checkout).  The xyz files next to this module are the
reference's own test inputs, copied unchanged from tests/data/nist-lj,
tests/data/nist-spce and tests/data/md-nacl.

Units: lumol internal units (A, fs, u, K; energy u*A^2/fs^2) unless stated.
"""

# lumol-core/src/consts.rs:9,15
K_BOLTZMANN = 8.31446284161522e-7
FOUR_PI_EPSILON_0 = 7.197589831304046
# lumol-core/src/units.rs:30-80
U_IN_KG = 1.660538782e-27
AVOGADRO = 6.02214179e23
KJ_PER_MOL = 1e-7 / U_IN_KG / AVOGADRO            # from(1, "kJ/mol")
KCAL_PER_MOL = 4.184 * 1e-7 / U_IN_KG / AVOGADRO   # from(1, "kcal/mol")
ATM = 101325.0 * 1e-40 / U_IN_KG                   # from(1, "atm")

# --- scalar functions ----------------------------------------------------------
# lumol-core/src/energy/functions.rs:622-644  (LennardJones{epsilon 0.8, sigma 2.0})
LJ_UNIT = {
    "sigma": 2.0, "epsilon": 0.8,
    "energy": [(2.0, 0.0, "exact"), (2.5, -0.6189584744448002, "exact")],
    "tail_energy": [(1.0, 1388.0888888888887), (2.0, -5.688888888888889),
                    (14.42, -0.022767318648783084)],
    "tail_virial": [(1.0, 17066.666666666668), (2.0, -17.06666666666667),
                    (14.42, -0.1366035877536718)],
    "force": [(2.5, -0.95773475733504, "ulps")],
}
# lumol-core/src/energy/functions.rs:62-67 (doc test, LJ sigma 2 eps 10)
LJ_DOC = {"sigma": 2.0, "epsilon": 10.0,
          "energy": [(2.0, 0.0), (3.0, -3.203365942785746)], "force": [(2.0, 120.0)]}
# lumol-core/src/energy/functions.rs:648-660 (Harmonic k 50, x0 2)
HARMONIC_UNIT = {"k": 50.0, "x0": 2.0,
                 "energy": [(2.0, 0.0), (2.5, 6.25)], "force": [(2.0, 0.0), (2.5, -25.0)]}
# lumol-core/src/energy/pairs.rs:335-360 (LJ sigma 1 eps 2, cutoff 4)
SHIFTED_PAIR = {"sigma": 1.0, "epsilon": 2.0, "cutoff": 4.0,
                "shifted_energy_2_5": -0.030681134109158216,   # ulps
                "tail_energy": -0.041663275824652776,          # exact
                "tail_virial_trace": -0.24995930989583334}     # ulps
# lumol-core/src/energy/pairs.rs:299-308 doc test (LJ sigma 0.5 eps 4.2 rc 2.0)
PAIR_TAIL_DOC = {"sigma": 0.5, "epsilon": 4.2, "cutoff": 2.0,
                 "tail_energy": -0.010936609903971353,
                 "tail_virial_diag": -0.02187143961588542}

# --- Wolf ------------------------------------------------------------------------------
# lumol-core/src/energy/global/wolf.rs:26-48: Na(+1) at 0, Cl(-1) at (2,0,0),
# cubic cell 30, Wolf(12): potential_energy() == -0.0729290269539354 (exact)
WOLF_DOC = {"cell": 30.0, "cutoff": 12.0, "energy": -0.0729290269539354}

# --- Ewald -----------------------------------------------------------------------------
# lumol-core/src/energy/global/ewald.rs:964-973, 1040-1055: Cl(-1) at 0, Na(+1)
# at (1.5,0,0), cell 20, Ewald(8, kmax 10): energy ~ -0.09262397663346732 (eps 1e-4)
EWALD_NACL_PAIR = {"cell": 20.0, "cutoff": 8.0, "kmax": 10,
                   "energy": -0.09262397663346732, "abs_tol": 1e-4}
# lumol-core/src/energy/global/ewald.rs:975-993, 1122-1134: water molecule,
# cell 20, Ewald(8, kmax 10), InterMolecular: energy == -9.243868813825495e-6 (ulps)
EWALD_WATER = {"cell": 20.0, "cutoff": 8.0, "kmax": 10,
               "positions": [(0.0, 0.0, 0.0), (-0.7, -0.7, 0.3), (0.3, -0.3, -0.8)],
               "names": ["O", "H", "H"], "charges": [-0.8476, 0.4238, 0.4238],
               "bonds": [(0, 1), (0, 2)],
               "energy": -0.000009243868813825495}
# lumol-core/src/energy/global/ewald.rs:1242-1335: two water molecules, cost of
# moving molecule 0 == E_new - E_old at max_relative 1e-12
EWALD_MOVE = {
    "cell": 20.0,
    "names": ["H", "O", "H", "H", "O", "H"],
    "positions": [(0.3, -0.3, -0.8), (0.0, 0.0, 0.0), (-0.7, -0.7, 0.3),
                  (2.3, 1.7, -0.8), (2.0, 2.0, 0.0), (1.3, 1.3, 0.3)],
    "bonds": [(0, 1), (1, 2), (3, 4), (4, 5)],
    "new_positions": [(0.41727, 2.29401, -0.0558),
                      (0.5097743599026461, 3.194114034722624, -0.020364564697826326),
                      (-0.2501317777731211, 3.562366060753896, -0.6178033542374419)],
}

# --- NIST LJ (tests/nist-lj.rs); systems tests/data/nist-lj/lj-{1..4}.xyz ----
# LJ epsilon 1 sigma 1 ("H","H"), no shift.  (energy, decimals), (trace W, decimals),
# (tail energy, decimals); assert_eq on values rounded at the given decimal.
NIST_LJ = {
    3.0: {  # tests/nist-lj.rs:58-129
        "lj-1.xyz": ((-4351.5, 1), (-568.67, 2), (-198.49, 2)),
        "lj-2.xyz": ((-690.00, 2), (-568.46, 2), (-24.230, 3)),
        "lj-3.xyz": ((-1146.7, 1), (-1164.9, 1), (-49.622, 3)),
        "lj-4.xyz": ((-16.790, 3), (-46.249, 3), (-0.54517, 5)),
    },
    4.0: {  # tests/nist-lj.rs:163-240
        "lj-1.xyz": ((-4467.5, 1), (-1263.9, 1), (-83.769, 3)),
        "lj-2.xyz": ((-704.60, 2), (-655.99, 2), (-10.226, 3)),
        "lj-3.xyz": ((-1175.4, 1), (-1337.1, 1), (-20.942, 3)),
        "lj-4.xyz": ((-17.060, 3), (-47.869, 3), (-0.23008, 5)),
    },
}

# --- NIST SPC/E (tests/nist-spce.rs) ------------------------------------------------
# O-O LJ eps 78.19743111 K*kB sigma 3.16555789 (+tail), O-H/H-H Null, Ewald
# alpha 5.6/L kmax 5 InterMolecular; qH 0.4238 qO -0.8476; bonds O-H1, O-H2.
# Values are energies / K_BOLTZMANN: (total, pairs, tail, coulomb), rel 1e-3.
NIST_SPCE_LJ = {"epsilon_K": 78.19743111, "sigma": 3.16555789, "qH": 0.42380}
NIST_SPCE_ENERGIES = {
    9.0: {  # tests/nist-spce.rs:136-299
        "spce-1.xyz": (-4.88608e5, 9.98560e4, -1.12959e3, -5.87334e5),
        "spce-2.xyz": (-1.06602e6, 1.94941e5, -4.51836e3, -1.25645e6),
        "spce-3.xyz": (-1.71488e6, 3.57106e5, -1.01663e4, -2.06205e6),
        "spce-4.xyz": (-3.08010e6, 4.53536e5, -1.88265e4, -3.51481e6),
    },
    10.0: {  # tests/nist-spce.rs:337-500
        "spce-1.xyz": (-4.88604e5, 9.95387e4, -8.23715e2, -5.87319e5),
        "spce-2.xyz": (-1.06590e6, 1.93712e5, -3.29486e3, -1.25632e6),
        "spce-3.xyz": (-1.71488e6, 3.54344e5, -7.41343e3, -2.06182e6),
        "spce-4.xyz": (-3.20501e6, 4.48593e5, -1.37286e4, -3.63987e6),
    },
}
# LAMMPS forces (kcal/mol/A), tests/nist-spce.rs:159-329, 360-531: (kmax, alpha,
# forces file); no tail; per-component relative tolerance 5e-3 (spce-1, spce-2) and
# the dynamic 1e-1 / 5e-2 / 1e-2 tolerance for spce-3 and spce-4 (:262-275, :318-329).
NIST_SPCE_LAMMPS = {
    9.0: {"spce-1.xyz": (8, 0.364209, "forces-9-1.xyz"),
          "spce-2.xyz": (8, 0.370036, "forces-9-2.xyz"),
          "spce-3.xyz": (8, 0.373403, "forces-9-3.xyz"),
          "spce-4.xyz": (12, 0.370914, "forces-9-4.xyz")},
    10.0: {"spce-1.xyz": (7, 0.326983, "forces-10-1.xyz"),
           "spce-2.xyz": (8, 0.332241, "forces-10-2.xyz"),
           "spce-3.xyz": (8, 0.335278, "forces-10-3.xyz"),
           "spce-4.xyz": (11, 0.333033, "forces-10-4.xyz")},
}
# Ewald components / K_BOLTZMANN (lumol-core/src/energy/global/ewald.rs:1470-1903):
# real (rel 1e-4), k-space (rel 2e-3; 5e-3 for spce-2), self (rel 1e-4);
# alpha = 5.6/L, kmax 5.
NIST_EWALD_COMPONENTS = {
    9.0: {"spce-1.xyz": (2.251086e6, 6.27009e3, -2.84469e6),
          "spce-2.xyz": (4.4269e6, 6.03495e3, -5.68938e6),
          "spce-3.xyz": (6.46678e6, 5.24461e3, -8.53407e6),
          "spce-4.xyz": (1.07011e7, 7.58785e3, -1.42235e7)},
    10.0: {"spce-1.xyz": (2.251101e6, 6.27009e3, -2.84469e6),
           "spce-2.xyz": (4.42703e6, 6.03495e3, -5.68938e6),
           "spce-3.xyz": (6.46701e6, 5.24461e3, -8.53407e6),
           "spce-4.xyz": (1.057604e7, 7.58785e3, -1.42235e7)},
}
# Ewald virials in atm * volume (ewald.rs:1494-1903): (kmax, alpha, real 3x3,
# kspace 3x3, tolerance on the real part); kspace rel 1e-3.
NIST_EWALD_VIRIALS = {
    9.0: {
        "spce-1.xyz": (8, 0.364209,
                       [[-3852.8846, 264.92131, 15.263331], [264.92131, -3778.4993, -108.79484],
                        [15.263331, -108.79484, -1885.924]],
                       [[-124.54878, 5.3873858, 15.093282], [5.3873858, -111.92329, -36.333371],
                        [15.093282, -36.333371, -248.12515]], 1e-3),
        "spce-2.xyz": (8, 0.370036,
                       [[-6388.8957, -158.39688, -344.71965], [-158.39688, -6988.116, 443.24826],
                        [-344.71965, 443.24826, -7169.9086]],
                       [[-296.86013, -10.296217, -3.0451399], [-10.296217, -250.90827, -7.9310376],
                        [-3.0451399, -7.9310376, -299.97078]], 1e-3),
        "spce-3.xyz": (8, 0.373403,
                       [[-10038.765, -1366.2912, 95.727168], [-1366.2912, -12683.884, -213.42417],
                        [95.727168, -213.42417, -11624.416]],
                       [[-255.20669, 18.680095, -37.1178], [18.680095, -275.89451, -5.1062841],
                        [-37.1178, -5.1062841, -232.9918]], 1e-3),
        "spce-4.xyz": (12, 0.370914,
                       [[-5586.0358, -136.58999, -30.075826], [-136.58999, -5846.9018, -121.54568],
                        [-30.075826, -121.54568, -4988.8238]],
                       [[-482.8217, -17.877365, 11.908942], [-17.877365, -465.33957, 3.5274012],
                        [11.908942, 3.5274012, -536.23714]], 5e-3),
    },
    10.0: {
        "spce-1.xyz": (7, 0.326983,
                       [[-3916.4836, 268.02381, 23.686442], [268.02381, -3839.2663, -118.84454],
                        [23.686442, -118.84454, -1945.1106]],
                       [[-61.298853, 2.0228732, 6.70617], [2.0228732, -51.134442, -26.2246],
                        [6.70617, -26.2246, -188.90856]], 1e-3),
        "spce-2.xyz": (8, 0.332241,
                       [[-6515.4904, -161.25038, -340.79656], [-161.25038, -7099.813, 440.80231],
                        [-340.79656, 440.80231, -7285.3392]],
                       [[-171.16013, -7.5200295, -6.7910143], [-7.5200295, -140.58412, -5.5476543],
                        [-6.7910143, -5.5476543, -185.27155]], 1e-3),
        "spce-3.xyz": (8, 0.335278,
                       [[-10159.552, -1365.6921, 85.837259], [-1365.6921, -12806.406, -202.68293],
                        [85.837259, -202.68293, -11744.811]],
                       [[-136.28835, 18.49649, -27.386553], [18.49649, -155.56804, -15.667645],
                        [-27.386553, -15.667645, -114.48139]], 1e-3),
        "spce-4.xyz": (11, 0.333033,
                       [[-5782.9498, -144.67187, -31.932323], [-144.67187, -6052.8821, -121.95614],
                        [-31.932323, -121.95614, -5205.1646]],
                       [[-288.2612, -9.676199, 13.775332], [-9.676199, -261.38081, 3.8066323],
                        [13.775332, 3.8066323, -322.02403]], 1e-3),
    },
}

# --- 2-atom compute tests (lumol-core/src/sys/compute.rs:537-560) -------------
# F at 0 and 1.3 in a cubic cell of 10, Harmonic{k 300 kJ/mol/A^2, x0 1.2}.
TWO_ATOM = {"cell": 10.0, "positions": [(0.0, 0.0, 0.0), (1.3, 0.0, 0.0)],
            "k_kj": 300.0, "x0": 1.2}

# --- MD NaCl small (tests/md-nacl.rs:17-31, tests/data/md-nacl) ----------------
# 64 ions, cubic 11.2804 A, LJ pairs (kcal/mol), Ewald(5.5, kmax 10), 100 VV
# steps of 1 fs from 300 K: |E_final - E_initial| / |E_final| < 5e-3 (Ewald).
MD_NACL_SMALL = {
    "cell": 11.2804, "cutoff": 5.5, "kmax": 10,
    "lj_kcal": {("Na", "Na"): (2.497, 0.07826), ("Cl", "Cl"): (4.612, 0.02502),
                ("Na", "Cl"): (3.5545, 0.04425)},
    "charges": {"Na": 1.0, "Cl": -1.0}, "masses": {"Na": 22.98976928, "Cl": 35.45},
    "drift_tol": 5e-3,
}

# --- other potentials (lumol-core/src/energy/functions.rs tests, doc tests) ----
# Each: (constructor parameters in the reference's order, [(what, x, value, check)])
# check: "eq" assert_eq!, "ulps" assert_ulps_eq!, ("abs", eps) |a - b| <= eps,
# ("rel", eps) approx::assert_relative_eq!(.., epsilon = eps).
OTHER_POTENTIALS = {
    # functions.rs:698-716
    "buckingham": ({"a": 2.0, "c": 1.0, "rho": 2.0}, [
        ("energy", 2.0, 0.7201338823428847, "eq"),
        ("force", 2.0, 0.32100444117144233, "eq"),
        ("tail_energy", 10.0, 1.8323882504179136, "eq"),
        ("tail_virial", 10.0, 33.422487868546725, "eq"),
    ]),
    # functions.rs:272-274 (doc test, Buckingham{a 2, c 1, rho 5.3})
    "buckingham_doc": ({"a": 2.0, "c": 1.0, "rho": 5.3}, [
        ("energy", 2.2, 1.3117360696239022, "eq"),
        ("force", 2.2, 0.2251072178835946, "eq"),
    ]),
    # functions.rs:718-738
    "born": ({"a": 2.0, "c": 1.0, "d": 0.5, "sigma": 2.0, "rho": 2.0}, [
        ("energy", 2.0, 1.986328125, "eq"),
        ("force", 2.0, 0.9609375, "eq"),
        ("tail_energy", 10.0, 4.981521444402363, "eq"),
        ("tail_virial", 10.0, 69.13986044386026, "eq"),
    ]),
    # functions.rs:335-337 (doc test)
    "born_doc": ({"a": 2.0, "c": 1.0, "d": 0.5, "sigma": 1.5, "rho": 5.3}, [
        ("energy", 2.2, 1.7446409593340713, "eq"),
        ("force", 2.2, 0.30992873382584607, "eq"),
    ]),
    # functions.rs:740-758
    "morse": ({"a": 2.0, "x0": 1.3, "depth": 4.0}, [
        ("energy", 1.0, 2.703517287822119, "eq"),
        ("force", 1.0, -37.12187076378477, "eq"),
        ("tail_energy", 1.0, 0.0, "eq"),
        ("tail_virial", 1.0, 0.0, "eq"),
    ]),
    # functions.rs:760-778
    "gaussian": ({"a": 8.0, "b": 2.0}, [
        ("energy", 0.0, -8.0, "eq"),
        ("force", 0.0, 0.0, "eq"),
        ("tail_energy", 2.5, -1.93518e-5, ("rel", 1e-10)),
        ("tail_virial", 2.5, 5.23887e-4, ("rel", 1e-10)),
    ]),
    # functions.rs:781-802
    "mie": ({"sigma": 2.0, "epsilon": 0.8, "n": 12.0, "m": 6.0}, [
        ("energy", 2.0, 0.0, "eq"),
        ("energy", 2.5, -0.6189584744448002, "eq"),
        ("tail_energy", 1.0, 1388.0888889, ("rel", 1e-6)),
        ("tail_energy", 2.0, -5.688888888888889, ("rel", 1e-6)),
        ("tail_energy", 14.42, -0.022767318648783084, ("rel", 1e-6)),
        ("tail_virial", 1.0, 17066.666666666668, ("rel", 1e-6)),
        ("tail_virial", 2.0, -17.06666666666667, ("rel", 1e-6)),
        ("tail_virial", 14.42, -0.1366035877536718, ("rel", 1e-6)),
        ("force", 2.5, -0.95773475733504, "ulps"),
    ]),
    # functions.rs:519-524 (doc test, Mie(sigma 2, epsilon 10, n 12, m 6))
    "mie_doc": ({"sigma": 2.0, "epsilon": 10.0, "n": 12.0, "m": 6.0}, [
        ("energy", 2.0, 0.0, "eq"),
        ("energy", 3.0, -3.203365942785746, ("abs", 1e-8)),
        ("force", 2.0, 120.0, "eq"),
    ]),
    # functions.rs:172-177 (doc test, CosineHarmonic::new(100, 2))
    "cosine_harmonic_doc": ({"k": 100.0, "x0": 2.0}, [
        ("energy", 2.0, 0.0, "eq"),
        ("energy", 3.0, 16.464942078100552, "eq"),
        ("force", 2.0, 0.0, "eq"),
    ]),
    # functions.rs:225-231 (doc test, Torsion{delta pi/2, k 10, n 3})
    "torsion_doc": ({"k": 10.0, "delta": 1.5707963267948966, "n": 3}, [
        ("energy", 1.5707963267948966, 0.0, "eq"),
        ("energy", 1.0471975511965976, 10.0, "eq"),
        ("force", 1.5707963267948966, 0.0, ("abs", 1e-12)),
    ]),
    # functions.rs:680-696 (Torsion{k 5, n 3, delta 3})
    "torsion": ({"k": 5.0, "delta": 3.0, "n": 3}, [
        ("energy", 1.0, 10.0, "eq"),
        ("force", 1.0, 0.0, "eq"),
    ]),
}

# lumol-core/src/energy/computations.rs:178-206 (Harmonic k 50 x0 2, 1000 points,
# max 4) and :94-101 (doc test: Harmonic x0 0.5 k 4.2, 1000 points, max 2).
# (x, energy, force), assert_eq!.
TABLE_HARMONIC = {"k": 50.0, "x0": 2.0, "size": 1000, "max": 4.0, "values": [
    (2.5, 6.25, -25.0),
    (4.0 - 2.0 * (4.0 / 1000.0), 99.2016, -99.6),
    (4.0 - 4.0 / 1000.0, 0.0, 0.0),
    (4.0, 0.0, 0.0),
    (4.1, 0.0, 0.0),
]}
TABLE_DOC = {"k": 4.2, "x0": 0.5, "size": 1000, "max": 2.0, "values": [(1.0, 0.525, None), (3.0, 0.0, None)]}
