// Compile-time constants of the full-band ensemble Monte Carlo path.
// Shared by the host library (g++/hipcc host pass) and the HIP kernels.
//
// Values follow include/Constant_GlobalVariable_MathFunction.h of the reference
// (table caps MTAB/MWLE/MWLES, band counts, motion-rule and direction tokens) and
// include/defs.h (MY_ZERO containment tolerance).
#pragma once

namespace emc {

// ---- mathematical / physical constants (Constant_GlobalVariable_MathFunction.h:15-79)
constexpr double PI = 3.141592653589793238462643e0;
constexpr double TWOPI = 2.0 * PI;
constexpr double EC = 1.602176e-19;       // electron charge [As]
constexpr double EM = 9.109383e-31;       // electron rest mass [kg]
constexpr double PLANCK = 6.582119e-16;   // hbar [eVs]
constexpr double CLIGHT = 299792458;      // [m/s]
constexpr double BOLTZ = 8.617343e-5;     // [eV/K]
constexpr double EPS0 = 8.854187817e-12;  // [As/Vm]

// ---- band structure sizes (Constant_GlobalVariable_MathFunction.h:83-121)
constexpr int NBE = 4;         // conduction bands
constexpr int NBH = 3;         // valence bands
constexpr int NBOE = 1;        // oxide conduction bands
constexpr int MNB = NBE + NBH + NBOE;
constexpr int MWLE = 325;      // coarse (40 meV) tetrahedron-list bins
constexpr int MWLES = 60;      // fine (2 meV) tetrahedron-list bins
constexpr int MTAB = 7500;     // 1 meV energy tables, entries 0..MTAB
constexpr int NTAB1 = MTAB + 1;
constexpr int MNScaEle = 14;   // electron processes (13 phonon + impact ionisation)
constexpr int MNScaHole = 4;   // hole processes (3 phonon + impact ionisation)
constexpr int NSYM = 48;

// ---- particle types (Constant_GlobalVariable_MathFunction.h:236)
constexpr int NumParType = 3, PELEC = 0, PHOLE = 1, POXEL = 2;

// ---- materials (Constant_GlobalVariable_MathFunction.h:173-185)
constexpr int VOID_MAT = 0, VACUUM = 1, OXIDE = 2, SILICON = 3, IGZO = 6;
constexpr int MNMaterialType = 7;

// ---- directions (Constant_GlobalVariable_MathFunction.h:203)
constexpr int UP = 0, DOWN = 1, LEFT = 2, RIGHT = 3, FRONT = 4, BACK = 5;

// ---- motion rules (Constant_GlobalVariable_MathFunction.h:211-227)
constexpr int PASS = 1, REFLECT = 2, SCATTOX = 3, PERIOD = 4, GENERATE = 5,
              CATCH = 6, CATCHGATE = 7, GENREF = 8;

constexpr int MNContact = 16;
constexpr int MNContactPlane = 10;
constexpr int MNSurface = 30;

// ---- tolerances (include/defs.h:126)
constexpr double MY_ZERO = 1e-13;

// Dispersive phonon final-energy tables (init_deep, src/ParticleUtils.cpp:750-808):
// kloem kloab ktoem ktoab klaem klaab, rows (final-energy classes) x columns (energy bins).
constexpr int PTAB_N = 6;
constexpr int PTAB_ROWS[PTAB_N] = {201, 201, 51, 51, 51, 51};
constexpr int PTAB_COLS[PTAB_N] = {120, 110, 160, 150, 160, 150};

// ---- packed particle word: type(1) | i(10) | j(11) | k(10); all ones = empty slot
constexpr unsigned PK_DEAD = 0xFFFFFFFFu;
constexpr int PK_IBITS = 10, PK_JBITS = 11, PK_KBITS = 10;
constexpr int PK_MAXI = (1 << PK_IBITS) - 1;
constexpr int PK_MAXJ = (1 << PK_JBITS) - 1;
constexpr int PK_MAXK = (1 << PK_KBITS) - 1;

}  // namespace emc
