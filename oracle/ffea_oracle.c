/*
 * ffea_oracle.c -- FP64 CPU restatement of FFEA's massless (CG_nomass) time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path in
 * ffea_amd/csrc: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links it and has no CPU
 * fallback through it.
 *
 * It restates, in plain C over flat arrays, the reference's algorithm in the
 * reference's own evaluation order (build with -O2 -ffp-contract=off, no
 * fast-math, so every sum and product rounds exactly as the reference's
 * strict-FP build does):
 *
 *   RngStream (MRG32k3a)       src/RngStream.cpp:45-105 (constants),
 *                              :112-217 (mod-m matrix ops), :266-310 (U01/U01d),
 *                              :327-435 (streams, substreams, AdvanceState),
 *                              :508-524 (RandU01/RandInt)
 *   package seeding            src/World.cpp:287-323
 *   element kernels            src/tetra_element_linear.cpp:193-205 (jacobian),
 *                              :223-293 (dpsi + volume + inversion),
 *                              :299-328 + :599-638 (viscosity matrix),
 *                              :378-408 (elastic stress), :417-436 (noise),
 *                              :441-449 + :474-490 (nodal forces),
 *                              :567-591 (midpoint re-linearisation)
 *                              include/tetra_element_linear.h:55-118 (macros)
 *   math helpers               src/mat_vec_fns.cpp:93-177, :250-275
 *   Blob glue                  src/Blob.cpp:335-412 (init order), :638-735
 *                              (update_internal_forces), :737-764
 *                              (update_positions), :1066-1085 (linearise_force),
 *                              :3760-3822 (rest state), :3826-3932
 *                              (aggregate_forces_and_solve), :3937-3953 (euler),
 *                              :3974-4019 (connectivity), :2861-2866 (material
 *                              unit conversion), :1424-1432 (strain energy)
 *   solver                     src/NoMassCGSolver.cpp:63-162 (init / pattern),
 *                              :165-197 (solve), :205-292 (vector kernels)
 *   sparse pattern / matrix    src/SparsityPattern.cpp:57-143,
 *                              src/SparseMatrixFixedPattern.cpp:48-68, :85-93,
 *                              :148-191, :325-333; src/SparseMatrixTypes.cpp:52-58
 *   step order                 src/World.cpp:1811-2043 (periodic box :1850-1901)
 *   mass-ful solvers           src/ConjugateGradientSolver.cpp:62-230,
 *                              src/MassLumpedSolver.cpp:40-90, src/Blob.cpp:1385-1407
 *                              (kinetic energy)
 *   forces outside elements    src/Blob.cpp:2094-2162 (CTF 1/2/3), :3350-3424
 *                              (surface forces); src/World.cpp:3808-3846 (springs)
 *   surface interactions       src/VdW_solver.cpp:31-73, :158-384, :635-661,
 *                              :713-784; src/LinkedListCube.tpp:282-299;
 *                              src/Face.cpp:197-261; src/Blob.cpp:3494-3578
 *
 * Layout is the reference's: all N_t nodes (linear + quadratic midpoints),
 * 10-node elements, a scalar CSR over 3*N_t rows whose every entry keeps the
 * list of element-matrix locations it sums.
 *
 * Parity pins (tests/): RngStream known-answer sum 39.697547445251
 * (tests/consistency/rngstream/testRngStream.cpp); sequential-vs-jump stream
 * equality; bitwise equality against the reference's own compiled objects
 * (oracle/_ref) for elements, assembled matrix, CG solutions (massless and
 * mass-ful), trajectories, and the rotational force (intersectingPointToLine);
 * closed forms (unit tetrahedron, rest state stress-free, rigid motions, the
 * momentum and energy sums of the interaction forces, LJ minimum at Rmin,
 * soft-core continuity at Rmin); the manual's ctforces example; the reference's
 * cubes_lj golden energy (-1.7870e-19 J within 1 %).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* RngStream restatement (MRG32k3a, L'Ecuyer et al. 2002)                      */
/* ------------------------------------------------------------------------- */

static const double m1 = 4294967087.0;
static const double m2 = 4294944443.0;
static const double norm_ = 1.0 / (4294967087.0 + 1.0);
static const double a12 = 1403580.0;
static const double a13n = 810728.0;
static const double a21 = 527612.0;
static const double a23n = 1370589.0;
static const double two17 = 131072.0;
static const double two53 = 9007199254740992.0;
static const double fact = 5.9604644775390625e-8;

static const double InvA1[3][3] = {{184888585.0, 0.0, 1945170933.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}};
static const double InvA2[3][3] = {{0.0, 360363334.0, 4225571728.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}};
static const double A1p0[3][3] = {{0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}, {-810728.0, 1403580.0, 0.0}};
static const double A2p0[3][3] = {{0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}, {-1370589.0, 0.0, 527612.0}};
static const double A1p76[3][3] = {{82758667.0, 1871391091.0, 4127413238.0},
                                   {3672831523.0, 69195019.0, 1871391091.0},
                                   {3672091415.0, 3528743235.0, 69195019.0}};
static const double A2p76[3][3] = {{1511326704.0, 3759209742.0, 1610795712.0},
                                   {4292754251.0, 1511326704.0, 3889917532.0},
                                   {3859662829.0, 4292754251.0, 3708466080.0}};
static const double A1p127[3][3] = {{2427906178.0, 3580155704.0, 949770784.0},
                                    {226153695.0, 1230515664.0, 3580155704.0},
                                    {1988835001.0, 986791581.0, 1230515664.0}};
static const double A2p127[3][3] = {{1464411153.0, 277697599.0, 1610723613.0},
                                    {32183930.0, 1464411153.0, 1022607788.0},
                                    {2824425944.0, 32183930.0, 2093834863.0}};

/* (a*s + c) mod m for |a|,|s|,|c| < 2^35, exact in doubles (src/RngStream.cpp:112-130). */
static double mult_mod_m(double a, double s, double c, double m) {
    double v = a * s + c;
    long a1;
    if (v >= two53 || v <= -two53) {
        a1 = (long)(a / two17);
        a -= a1 * two17;
        v = a1 * s;
        a1 = (long)(v / m);
        v -= a1 * m;
        v = v * two17 + a * s + c;
    }
    a1 = (long)(v / m);
    if ((v -= a1 * m) < 0.0) return v += m;
    return v;
}

static void mat_vec_mod_m(const double A[3][3], const double s[3], double v[3], double m) {
    double x[3];
    for (int i = 0; i < 3; ++i) {
        x[i] = mult_mod_m(A[i][0], s[0], 0.0, m);
        x[i] = mult_mod_m(A[i][1], s[1], x[i], m);
        x[i] = mult_mod_m(A[i][2], s[2], x[i], m);
    }
    for (int i = 0; i < 3; ++i) v[i] = x[i];
}

static void mat_mat_mod_m(const double A[3][3], const double B[3][3], double C[3][3], double m) {
    double V[3], W[3][3];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) V[j] = B[j][i];
        mat_vec_mod_m(A, V, V, m);
        for (int j = 0; j < 3; ++j) W[j][i] = V[j];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[i][j] = W[i][j];
}

static void mat_two_pow_mod_m(const double A[3][3], double B[3][3], double m, long e) {
    if (A != (const double(*)[3])B)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) B[i][j] = A[i][j];
    for (long i = 0; i < e; i++) mat_mat_mod_m((const double(*)[3])B, (const double(*)[3])B, B, m);
}

static void mat_pow_mod_m(const double A[3][3], double B[3][3], double m, long n) {
    double W[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            W[i][j] = A[i][j];
            B[i][j] = 0.0;
        }
    for (int j = 0; j < 3; ++j) B[j][j] = 1.0;
    while (n > 0) {
        if (n % 2) mat_mat_mod_m((const double(*)[3])W, (const double(*)[3])B, B, m);
        mat_mat_mod_m((const double(*)[3])W, (const double(*)[3])W, W, m);
        n /= 2;
    }
}

typedef struct {
    double Cg[6], Bg[6], Ig[6];
    int anti, inc_prec;
} orc_stream;

static double next_seed[6] = {12345.0, 12345.0, 12345.0, 12345.0, 12345.0, 12345.0};

static int check_seed(const unsigned long seed[6]) {
    for (int i = 0; i < 3; ++i)
        if (seed[i] >= 4294967087UL) return -1;
    for (int i = 3; i < 6; ++i)
        if (seed[i] >= 4294944443UL) return -1;
    if (seed[0] == 0 && seed[1] == 0 && seed[2] == 0) return -1;
    if (seed[3] == 0 && seed[4] == 0 && seed[5] == 0) return -1;
    return 0;
}

/* RngStream::SetPackageSeed (src/RngStream.cpp:378-385); returns 1 on success. */
int orc_rng_set_package_seed(const unsigned long seed[6]) {
    if (check_seed(seed)) return 0;
    for (int i = 0; i < 6; ++i) next_seed[i] = (double)seed[i];
    return 1;
}

/* Constructor RngStream::RngStream (src/RngStream.cpp:327-345). */
void orc_rng_stream_init(orc_stream *g) {
    g->anti = 0;
    g->inc_prec = 0;
    for (int i = 0; i < 6; ++i) g->Bg[i] = g->Cg[i] = g->Ig[i] = next_seed[i];
    mat_vec_mod_m(A1p127, next_seed, next_seed, m1);
    mat_vec_mod_m(A2p127, &next_seed[3], &next_seed[3], m2);
}

/* RngStream::U01 (src/RngStream.cpp:266-289). */
double orc_rng_u01(orc_stream *g) {
    long k;
    double p1, p2, u;
    p1 = a12 * g->Cg[1] - a13n * g->Cg[0];
    k = (long)(p1 / m1);
    p1 -= k * m1;
    if (p1 < 0.0) p1 += m1;
    g->Cg[0] = g->Cg[1];
    g->Cg[1] = g->Cg[2];
    g->Cg[2] = p1;
    p2 = a21 * g->Cg[5] - a23n * g->Cg[3];
    k = (long)(p2 / m2);
    p2 -= k * m2;
    if (p2 < 0.0) p2 += m2;
    g->Cg[3] = g->Cg[4];
    g->Cg[4] = g->Cg[5];
    g->Cg[5] = p2;
    u = ((p1 > p2) ? (p1 - p2) * norm_ : (p1 - p2 + m1) * norm_);
    return (g->anti == 0) ? u : (1 - u);
}

static double orc_rng_u01d(orc_stream *g) {
    double u = orc_rng_u01(g);
    if (g->anti) {
        u += (orc_rng_u01(g) - 1.0) * fact;
        return (u < 0.0) ? u + 1.0 : u;
    }
    u += orc_rng_u01(g) * fact;
    return (u < 1.0) ? u : (u - 1.0);
}

double orc_rng_randu01(orc_stream *g) { return g->inc_prec ? orc_rng_u01d(g) : orc_rng_u01(g); }

int orc_rng_randint(orc_stream *g, int low, int high) {
    return low + (int)((high - low + 1.0) * orc_rng_randu01(g));
}

void orc_rng_reset_start_stream(orc_stream *g) {
    for (int i = 0; i < 6; ++i) g->Cg[i] = g->Bg[i] = g->Ig[i];
}
void orc_rng_reset_start_substream(orc_stream *g) {
    for (int i = 0; i < 6; ++i) g->Cg[i] = g->Bg[i];
}
void orc_rng_reset_next_substream(orc_stream *g) {
    mat_vec_mod_m(A1p76, g->Bg, g->Bg, m1);
    mat_vec_mod_m(A2p76, &g->Bg[3], &g->Bg[3], m2);
    for (int i = 0; i < 6; ++i) g->Cg[i] = g->Bg[i];
}
void orc_rng_set_antithetic(orc_stream *g, int a) { g->anti = a != 0; }
void orc_rng_increased_precis(orc_stream *g, int p) { g->inc_prec = p != 0; }

int orc_rng_set_seed(orc_stream *g, const unsigned long seed[6]) {
    if (check_seed(seed)) return 0;
    for (int i = 0; i < 6; ++i) g->Cg[i] = g->Bg[i] = g->Ig[i] = (double)seed[i];
    return 1;
}

void orc_rng_get_state(const orc_stream *g, unsigned long seed[6]) {
    for (int i = 0; i < 6; ++i) seed[i] = (unsigned long)g->Cg[i];
}

/* RngStream::AdvanceState (src/RngStream.cpp:408-435): n = 2^e + c (e>0), -2^-e + c (e<0), c (e=0). */
void orc_rng_advance_state(orc_stream *g, long e, long c) {
    double B1[3][3], C1[3][3], B2[3][3], C2[3][3];
    if (e > 0) {
        mat_two_pow_mod_m(A1p0, B1, m1, e);
        mat_two_pow_mod_m(A2p0, B2, m2, e);
    } else if (e < 0) {
        mat_two_pow_mod_m(InvA1, B1, m1, -e);
        mat_two_pow_mod_m(InvA2, B2, m2, -e);
    }
    if (c >= 0) {
        mat_pow_mod_m(A1p0, C1, m1, c);
        mat_pow_mod_m(A2p0, C2, m2, c);
    } else {
        mat_pow_mod_m(InvA1, C1, m1, -c);
        mat_pow_mod_m(InvA2, C2, m2, -c);
    }
    if (e) {
        mat_mat_mod_m((const double(*)[3])B1, (const double(*)[3])C1, C1, m1);
        mat_mat_mod_m((const double(*)[3])B2, (const double(*)[3])C2, C2, m2);
    }
    mat_vec_mod_m((const double(*)[3])C1, g->Cg, g->Cg, m1);
    mat_vec_mod_m((const double(*)[3])C2, &g->Cg[3], &g->Cg[3], m2);
}

/* World::init RNG.2 (src/World.cpp:299-305): six seeds from glibc srand/rand.  The int
 * sum rand()+rand() may exceed INT_MAX (undefined in C++); the reference's optimised build
 * adds in the 64-bit type of twoMax, so the sum does not wrap.  The reference's checkpoint
 * tests/physics/cyl_160_fine/cyl_160_fine.fcp (rng_seed 753) holds the non-wrapping seeds. */
void orc_package_seed_from_rng_seed(int rng_seed, unsigned long sixseed[6]) {
    const unsigned long two_max[2] = {4294967087UL, 4294944443UL};
    srand((unsigned)rng_seed);
    for (int i = 0; i < 6; i++) {
        unsigned long r1 = (unsigned long)rand();
        unsigned long r2 = (unsigned long)rand();
        sixseed[i] = (r1 + r2) % two_max[i / 3];
    }
}

/* ------------------------------------------------------------------------- */
/* Element kernels (src/tetra_element_linear.cpp)                              */
/* ------------------------------------------------------------------------- */

typedef double m3[3][3];

/* calculate_jacobian (:193-205): J[r][c] = pos(n_{r+1})[c] - pos(n_0)[c]. */
static void elem_jacobian(const double *pos, const int *n, m3 J) {
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) J[r][c] = pos[3 * n[r + 1] + c] - pos[3 * n[0] + c];
}

/* calc_shape_function_derivatives_and_volume (:223-293).  Returns -1 on a sign change
 * of det relative to *last_det, leaving dpsi half-built exactly as the reference does. */
static int elem_dpsi_vol(m3 J, double dpsi[12], double *vol, double *last_det) {
    double det;
    dpsi[1] = J[2][2] * J[1][1] - J[2][1] * J[1][2];
    dpsi[2] = J[2][1] * J[0][2] - J[2][2] * J[0][1];
    dpsi[3] = J[1][2] * J[0][1] - J[1][1] * J[0][2];
    dpsi[5] = J[2][0] * J[1][2] - J[2][2] * J[1][0];
    dpsi[6] = J[2][2] * J[0][0] - J[2][0] * J[0][2];
    dpsi[7] = J[1][0] * J[0][2] - J[1][2] * J[0][0];
    dpsi[9] = J[2][1] * J[1][0] - J[2][0] * J[1][1];
    dpsi[10] = J[2][0] * J[0][1] - J[2][1] * J[0][0];
    dpsi[11] = J[1][1] * J[0][0] - J[1][0] * J[0][1];
    det = J[0][0] * dpsi[1] + J[1][0] * dpsi[2] + J[2][0] * dpsi[3];
    if (*last_det * det < 0) return -1;
    *last_det = det;
    *vol = (1.0 / 6.0) * fabs(det);
    det = 1.0 / det;
    dpsi[1] *= det;
    dpsi[2] *= det;
    dpsi[3] *= det;
    dpsi[5] *= det;
    dpsi[6] *= det;
    dpsi[7] *= det;
    dpsi[9] *= det;
    dpsi[10] *= det;
    dpsi[11] *= det;
    dpsi[0] = -(dpsi[1] + dpsi[2] + dpsi[3]);
    dpsi[4] = -(dpsi[5] + dpsi[6] + dpsi[7]);
    dpsi[8] = -(dpsi[9] + dpsi[10] + dpsi[11]);
    return 0;
}

/* create_viscosity_matrix (:299-328) with BULK_VISCOUS_SUBMATRIX_* (header :55-116),
 * calc_del2_matrix (:599-619) and add_diffusion_matrix (:621-638). */
static void elem_viscosity(const double dpsi[12], double A, double B, double vol, double V[12][12]) {
    /* diagonal 4x4 submatrices, upper triangle */
    for (int I = 0; I < 12; I += 4) {
        for (int a = 0; a < 4; a++)
            for (int b = a; b < 4; b++) V[I + a][I + b] = dpsi[I + a] * dpsi[I + b] * (A + B);
    }
    /* off-diagonal submatrices (I,J) = (0,4), (0,8), (4,8) */
    const int IJ[3][2] = {{0, 4}, {0, 8}, {4, 8}};
    for (int s = 0; s < 3; s++) {
        int I = IJ[s][0], J = IJ[s][1];
        double K[4][4];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) K[a][b] = dpsi[I + a] * dpsi[J + b];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++)
                V[I + a][J + b] = (a == b) ? (B + A) * K[a][a] : B * K[a][b] + A * K[b][a];
    }
    /* del2 (upper triangle) and diffusion term on the three diagonal blocks */
    double del2[4][4];
    for (int a = 0; a < 4; a++)
        for (int b = a; b < 4; b++)
            del2[a][b] = (dpsi[a] * dpsi[b] + dpsi[4 + a] * dpsi[4 + b] + dpsi[8 + a] * dpsi[8 + b]);
    for (int i = 0; i < 3; i++)
        for (int a = 0; a < 4; a++)
            for (int b = a; b < 4; b++) V[i * 4 + a][i * 4 + b] += del2[a][b] * A;
    for (int i = 0; i < 12; i++)
        for (int j = 0; j <= i; j++) V[j][i] *= vol;
    for (int i = 1; i < 12; i++)
        for (int j = 0; j < i; j++) V[i][j] = V[j][i];
}

/* add_shear_elastic_stress (:378-396) + mat3_mult_both_transposed / mat3_mult_transpose
 * / mat3_scale (src/mat_vec_fns.cpp:93-126).  stress must be zero on entry. */
static void elem_shear_stress(m3 J, const double Jinv0[9], double G, double vol_0, double vol, m3 F,
                              m3 stress) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) F[i][j] = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) F[i][j] += J[k][i] * Jinv0[3 * j + k];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) stress[i][j] += F[i][k] * F[j][k];
    double s = (G * vol_0 / vol);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) stress[i][j] *= s;
    stress[0][0] -= G;
    stress[1][1] -= G;
    stress[2][2] -= G;
}

/* add_bulk_elastic_stress (:401-408). */
static void elem_bulk_stress(double G, double E, double vol_0, double vol, m3 stress) {
    double c_2 = E - G * 2.0 / 3.0;
    double c = G * (1.0 - (vol_0 / vol)) + 0.5 * c_2 * ((vol / vol_0) - (vol_0 / vol));
    stress[0][0] += c;
    stress[1][1] += c;
    stress[2][2] += c;
}

/* RAND(-.5,.5) (include/tetra_element_linear.h:118). */
static double rand_half(orc_stream *g) { return ((-.5) + ((.5) - (-.5)) * (orc_rng_randu01(g))); }

/* add_fluctuating_stress (:417-436): 7 draws, order bf, s00, s11, s22, s01, s02, s12. */
static void elem_fluct_stress(double kT, double dt, double A, double B, double vol, orc_stream *g, m3 stress) {
    double c = sqrt((24 * kT) / (vol * dt));
    double bf = sqrt(B) * rand_half(g);
    stress[0][0] += c * (sqrt(2 * A) * rand_half(g) + bf);
    stress[1][1] += c * (sqrt(2 * A) * rand_half(g) + bf);
    stress[2][2] += c * (sqrt(2 * A) * rand_half(g) + bf);
    stress[0][1] += c * sqrt(A) * rand_half(g);
    stress[0][2] += c * sqrt(A) * rand_half(g);
    stress[1][2] += c * sqrt(A) * rand_half(g);
    stress[1][0] = stress[0][1];
    stress[2][0] = stress[0][2];
    stress[2][1] = stress[1][2];
}

/* linearise_element (:567-591). */
static void elem_linearise(double *pos, const int *n) {
    const int ends[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (int m = 0; m < 6; m++)
        for (int c = 0; c < 3; c++)
            pos[3 * n[4 + m] + c] = .5 * (pos[3 * n[ends[m][0]] + c] + pos[3 * n[ends[m][1]] + c]);
}

/* mat3_invert (src/mat_vec_fns.cpp:148-177). */
static void mat3_invert(m3 m, double inv[9]) {
    double mi[3][3], det;
    mi[0][0] = m[2][2] * m[1][1] - m[2][1] * m[1][2];
    mi[0][1] = m[2][1] * m[0][2] - m[2][2] * m[0][1];
    mi[0][2] = m[1][2] * m[0][1] - m[1][1] * m[0][2];
    mi[1][0] = m[2][0] * m[1][2] - m[2][2] * m[1][0];
    mi[1][1] = m[2][2] * m[0][0] - m[2][0] * m[0][2];
    mi[1][2] = m[1][0] * m[0][2] - m[1][2] * m[0][0];
    mi[2][0] = m[2][1] * m[1][0] - m[2][0] * m[1][1];
    mi[2][1] = m[2][0] * m[0][1] - m[2][1] * m[0][0];
    mi[2][2] = m[1][1] * m[0][0] - m[1][0] * m[0][1];
    det = m[0][0] * mi[0][0] + m[1][0] * mi[0][1] + m[2][0] * mi[0][2];
    det = 1.0 / det;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) inv[3 * i + j] = mi[i][j] * det;
}

/* ------------------------------------------------------------------------- */
/* Blob (N_t layout) + NoMassCGSolver restatement                              */
/* ------------------------------------------------------------------------- */

typedef struct {
    int col;
    int nsrc, cap;
    const double **src;
} orc_entry;

typedef struct {
    int n, cap;
    orc_entry *e;
} orc_row;

typedef struct {
    /* sizes */
    int num_nodes, num_elements, num_pinned;
    /* parameters in meso units (as held by SimulationParams after assign()) */
    double dt, kT, epsilon2, stokes_visc;
    int max_iterations_cg, calc_noise, calc_stokes;
    /* nodes */
    double *pos, *vel, *stokes_radius, *stokes_drag, *force;
    int *is_pinned, *pinned;
    /* elements */
    int *topo;
    double *rho, *A, *B, *G, *E, *mass;
    double *J_inv_0, *vol_0, *vol, *last_det, *dpsi, *F_ij, *visc, *node_force, *stress_mag;
    /* node -> element force contributions (elem*10 + slot) */
    int *contrib_off, *contrib;
    /* fixed-pattern CSR with source lists */
    int num_rows, nnz;
    int *key, *col, *src_off;
    const double **src;
    double *val, *diag_inv;
    int *diag_idx;
    double one;
    /* CG work vectors (vector3 per node in the reference) */
    double *r, *p, *z, *q, *f;
    int last_cg_iterations, last_cg_status, last_inversions;
    /* solver choice (include/FFEA_return_codes.h:30-33): 1 ConjugateGradientSolver on the
     * mass matrix, 2 MassLumpedSolver, 3 NoMassCGSolver (the default) */
    int linear_solver;
    int *mkey, *mcol;      /* mass matrix, scalar CSR over N_t rows */
    double *mval, *mprec;  /* its values and Jacobi preconditioner */
    double *inv_M;         /* lumped inverse masses [N_t] */
    /* constant forces (Blob::apply_ctforces, src/Blob.cpp:2094-2162): linear (CTF 1) and
     * linear surface (CTF 3) */
    int n_ctf;
    int *ctf_node;
    double *ctf_force;     /* [3 n_ctf] */
    int n_rctf;            /* rotational (CTF 2), point + direction axes only */
    int *rctf_node;
    double *rctf_F;        /* [n_rctf] meso force */
    double *rctf_axis;     /* [6 n_rctf] axis point (meso length), unit direction */
    int *rctf_torque;      /* [n_rctf] 0: 'f' (circular force), 1: 't' (torque) */
    int n_faces, n_sets;
    int *face_nodes;       /* [3 n_faces] */
    int *set_off, *set_faces;
    double *set_force;     /* [3 n_sets] */
    double *face_force;    /* [12 n_faces]: Face::force[0..3], the three corners then the opposite vertex */
    /* surface-surface interactions: per surface face its type (-1: inactive) and the
     * geometry Face::calc_area_normal_centroid last left (centroid, normal, area) */
    int *ssint_type;       /* [n_faces] or NULL */
    int *ssint_opp;        /* [n_faces] the face element's vertex opposite the face (Face::n[3]) */
    double *face_geom;     /* [7 n_faces] */
} orc_blob;

static void *xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s);
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}

/* SparsityPattern::register_contribution (src/SparsityPattern.cpp:57-80). */
static void pattern_register(orc_row *rows, int i, int j, const double *loc) {
    orc_row *R = &rows[i];
    int k;
    for (k = 0; k < R->n; k++) {
        if (R->e[k].col == j) {
            orc_entry *E = &R->e[k];
            if (E->nsrc == E->cap) {
                E->cap = E->cap ? 2 * E->cap : 4;
                E->src = realloc(E->src, sizeof(*E->src) * E->cap);
            }
            E->src[E->nsrc++] = loc;
            return;
        }
        if (R->e[k].col > j) break;
    }
    if (R->n == R->cap) {
        R->cap = R->cap ? 2 * R->cap : 8;
        R->e = realloc(R->e, sizeof(*R->e) * R->cap);
    }
    memmove(&R->e[k + 1], &R->e[k], sizeof(*R->e) * (R->n - k));
    R->n++;
    R->e[k].col = j;
    R->e[k].cap = 4;
    R->e[k].nsrc = 1;
    R->e[k].src = malloc(sizeof(*R->e[k].src) * 4);
    R->e[k].src[0] = loc;
}

static int pattern_check(const orc_row *rows, int i, int j) {
    for (int k = 0; k < rows[i].n; k++)
        if (rows[i].e[k].col == j) return 1;
    return 0;
}

/* Create a blob.  pos: [3*num_nodes] meso units (already scaled); topo: [10*E];
 * mat_si: [6*E] density, shear_visc, bulk_visc, shear_mod, bulk_mod, dielectric in SI;
 * stokes_radius: [num_nodes] meso units; params in meso units. */
orc_blob *orc_blob_create(int num_nodes, int num_elements, const double *pos, const int *topo,
                          const double *mat_si, const double *stokes_radius, int num_pinned,
                          const int *pinned, double dt, double kT, double epsilon2, int max_iterations_cg,
                          int calc_noise, int calc_stokes, double stokes_visc) {
    /* mesoDimensions (include/dimensions.h:30-41) */
    const double meso_volume = 4.913e-30, meso_mass = 1.994307387553024e-26;
    const double meso_pressure = 8.4305832688784826e8, meso_time = 3.7302915560886126e-13;
    const double pi = 3.141592653589793238462643383279502884197169399375105820974944592307816406286;

    orc_blob *b = xcalloc(1, sizeof(orc_blob));
    int N = num_nodes, E = num_elements;
    b->num_nodes = N;
    b->num_elements = E;
    b->dt = dt;
    b->kT = kT;
    b->epsilon2 = epsilon2;
    b->max_iterations_cg = max_iterations_cg;
    b->calc_noise = calc_noise;
    b->calc_stokes = calc_stokes;
    b->stokes_visc = stokes_visc;
    b->linear_solver = 3;
    b->pos = xcalloc(3 * (size_t)N, sizeof(double));
    b->vel = xcalloc(3 * (size_t)N, sizeof(double));
    b->force = xcalloc(3 * (size_t)N, sizeof(double));
    b->stokes_radius = xcalloc(N, sizeof(double));
    b->stokes_drag = xcalloc(N, sizeof(double));
    b->is_pinned = xcalloc(N, sizeof(int));
    memcpy(b->pos, pos, sizeof(double) * 3 * (size_t)N);
    memcpy(b->stokes_radius, stokes_radius, sizeof(double) * N);
    b->num_pinned = num_pinned;
    b->pinned = xcalloc(num_pinned, sizeof(int));
    for (int i = 0; i < num_pinned; i++) {
        b->pinned[i] = pinned[i];
        b->is_pinned[pinned[i]] = 1;
    }
    b->topo = xcalloc(10 * (size_t)E, sizeof(int));
    memcpy(b->topo, topo, sizeof(int) * 10 * (size_t)E);
    b->rho = xcalloc(E, sizeof(double));
    b->A = xcalloc(E, sizeof(double));
    b->B = xcalloc(E, sizeof(double));
    b->G = xcalloc(E, sizeof(double));
    b->E = xcalloc(E, sizeof(double));
    b->mass = xcalloc(E, sizeof(double));
    /* Blob::load_material_params unit conversion (src/Blob.cpp:2861-2866); the
     * (2/3)*shear_visc term is subtracted in SI units, as the reference does. */
    for (int e = 0; e < E; e++) {
        const double *m = &mat_si[6 * e];
        b->rho[e] = m[0] * meso_volume / meso_mass;
        b->A[e] = m[1] / (meso_pressure * meso_time);
        b->B[e] = m[2] / (meso_pressure * meso_time) - (2.0 / 3.0) * m[1];
        b->G[e] = m[3] / meso_pressure;
        b->E[e] = m[4] / meso_pressure;
    }
    b->J_inv_0 = xcalloc(9 * (size_t)E, sizeof(double));
    b->vol_0 = xcalloc(E, sizeof(double));
    b->vol = xcalloc(E, sizeof(double));
    b->last_det = xcalloc(E, sizeof(double));
    b->dpsi = xcalloc(12 * (size_t)E, sizeof(double));
    b->F_ij = xcalloc(9 * (size_t)E, sizeof(double));
    b->visc = xcalloc(144 * (size_t)E, sizeof(double));
    b->node_force = xcalloc(30 * (size_t)E, sizeof(double));
    b->stress_mag = xcalloc(E, sizeof(double));

    /* Blob::init order (src/Blob.cpp:335-412): linearise, rest state, connectivity,
     * solver init, then stokes drag. */
    for (int e = 0; e < E; e++) elem_linearise(b->pos, &b->topo[10 * e]);

    /* calc_rest_state_info (src/Blob.cpp:3760-3822) */
    for (int e = 0; e < E; e++) {
        m3 J;
        elem_jacobian(b->pos, &b->topo[10 * e], J);
        mat3_invert(J, &b->J_inv_0[9 * e]);
        elem_dpsi_vol(J, &b->dpsi[12 * e], &b->vol[e], &b->last_det[e]);
        b->vol_0[e] = b->vol[e];
        b->mass[e] = b->vol_0[e] * b->rho[e];
    }

    /* calculate_node_element_connectivity (src/Blob.cpp:3974-4019) */
    b->contrib_off = xcalloc((size_t)N + 1, sizeof(int));
    for (int e = 0; e < E; e++)
        for (int j = 0; j < 10; j++) b->contrib_off[b->topo[10 * e + j] + 1]++;
    for (int i = 0; i < N; i++) b->contrib_off[i + 1] += b->contrib_off[i];
    b->contrib = xcalloc((size_t)b->contrib_off[N], sizeof(int));
    {
        int *cnt = xcalloc(N, sizeof(int));
        for (int e = 0; e < E; e++)
            for (int j = 0; j < 10; j++) {
                int ni = b->topo[10 * e + j];
                b->contrib[b->contrib_off[ni] + cnt[ni]++] = 10 * e + j;
            }
        free(cnt);
    }

    /* NoMassCGSolver::init (src/NoMassCGSolver.cpp:63-162) */
    b->num_rows = 3 * N;
    b->one = 1;
    orc_row *rows = xcalloc(b->num_rows, sizeof(orc_row));
    for (int e = 0; e < E; e++) {
        m3 J;
        elem_jacobian(b->pos, &b->topo[10 * e], J);
        elem_dpsi_vol(J, &b->dpsi[12 * e], &b->vol[e], &b->last_det[e]);
        elem_viscosity(&b->dpsi[12 * e], b->A[e], b->B[e], b->vol[e], (double(*)[12]) & b->visc[144 * e]);
        for (int ni = 0; ni < 10; ++ni)
            for (int nj = 0; nj < 10; ++nj) {
                int ni_index = b->topo[10 * e + ni], nj_index = b->topo[10 * e + nj];
                int ni_row = 3 * ni_index, nj_row = 3 * nj_index;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        if (b->is_pinned[ni_index] == 0 && b->is_pinned[nj_index] == 0 && ni < 4 && nj < 4) {
                            pattern_register(rows, ni_row + i, nj_row + j,
                                             &b->visc[144 * e + 12 * (ni + 4 * i) + (nj + 4 * j)]);
                        } else if (ni == nj && i == j) {
                            if (!pattern_check(rows, ni_row + i, nj_row + j))
                                pattern_register(rows, ni_row + i, nj_row + j, &b->one);
                        }
                    }
            }
    }
    if (calc_stokes == 1)
        for (int ni = 0; ni < N; ++ni)
            if (b->is_pinned[ni] == 0)
                for (int nj = 0; nj < 3; ++nj) pattern_register(rows, 3 * ni + nj, 3 * ni + nj, &b->stokes_drag[ni]);

    /* SparsityPattern::create_sparse_matrix (src/SparsityPattern.cpp:97-143) */
    int nnz = 0;
    size_t nsrc = 0;
    for (int i = 0; i < b->num_rows; i++) {
        nnz += rows[i].n;
        for (int k = 0; k < rows[i].n; k++) nsrc += rows[i].e[k].nsrc;
    }
    b->nnz = nnz;
    b->key = xcalloc((size_t)b->num_rows + 1, sizeof(int));
    b->col = xcalloc(nnz, sizeof(int));
    b->val = xcalloc(nnz, sizeof(double));
    b->src_off = xcalloc((size_t)nnz + 1, sizeof(int));
    b->src = xcalloc(nsrc, sizeof(double *));
    b->diag_idx = xcalloc(b->num_rows, sizeof(int));
    int pos_ = 0;
    size_t sp = 0;
    for (int i = 0; i < b->num_rows; i++) {
        b->key[i] = pos_;
        for (int k = 0; k < rows[i].n; k++) {
            b->col[pos_] = rows[i].e[k].col;
            b->src_off[pos_] = (int)sp;
            for (int s = 0; s < rows[i].e[k].nsrc; s++) b->src[sp++] = rows[i].e[k].src[s];
            if (b->col[pos_] == i) b->diag_idx[i] = pos_;
            free(rows[i].e[k].src);
            pos_++;
        }
        free(rows[i].e);
    }
    b->key[b->num_rows] = pos_;
    b->src_off[nnz] = (int)sp;
    free(rows);
    b->diag_inv = xcalloc(b->num_rows, sizeof(double));
    b->r = xcalloc(3 * (size_t)N, sizeof(double));
    b->p = xcalloc(3 * (size_t)N, sizeof(double));
    b->z = xcalloc(3 * (size_t)N, sizeof(double));
    b->q = xcalloc(3 * (size_t)N, sizeof(double));
    b->f = xcalloc(3 * (size_t)N, sizeof(double));

    /* stokes drag (src/Blob.cpp:408-412) */
    if (calc_stokes == 1)
        for (int i = 0; i < N; ++i) b->stokes_drag[i] = 6.0 * pi * stokes_visc * b->stokes_radius[i];
    return b;
}

void orc_blob_destroy(orc_blob *b) {
    if (!b) return;
    void *ptrs[] = {b->pos, b->vel, b->stokes_radius, b->stokes_drag, b->force, b->is_pinned, b->pinned,
                    b->topo, b->rho, b->A, b->B, b->G, b->E, b->mass, b->J_inv_0, b->vol_0, b->vol,
                    b->last_det, b->dpsi, b->F_ij, b->visc, b->node_force, b->stress_mag, b->contrib_off,
                    b->contrib, b->key, b->col, b->src_off, (void *)b->src, b->val, b->diag_inv, b->diag_idx,
                    b->r, b->p, b->z, b->q, b->f, b->mkey, b->mcol, b->mval, b->mprec, b->inv_M,
                    b->ctf_node, b->ctf_force, b->rctf_node, b->rctf_F, b->rctf_axis, b->rctf_torque, b->face_nodes, b->set_off, b->set_faces, b->set_force,
                    b->face_force, b->ssint_type, b->ssint_opp, b->face_geom};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    free(b);
}

/* Overwrite all N_t node positions (the restart frame of src/World.cpp:647-763). */
void orc_blob_set_positions(orc_blob *b, const double *pos) {
    memcpy(b->pos, pos, sizeof(double) * 3 * (size_t)b->num_nodes);
}

/* Blob::zero_force + set_forces_to_zero (src/World.cpp:1844-1909). */
static void blob_zero_forces(orc_blob *b) {
    memset(b->node_force, 0, sizeof(double) * 30 * (size_t)b->num_elements);
    memset(b->force, 0, sizeof(double) * 3 * (size_t)b->num_nodes);
    if (b->face_force) memset(b->face_force, 0, sizeof(double) * 12 * (size_t)b->n_faces);
}

/* ---- forces from outside the element loop ----------------------------------- */

/* Linear constant forces, [3n] meso units, in the order Blob::load_ctforces stores them
 * (src/Blob.cpp:3184-3241).  n = 0 clears. */
void orc_blob_set_constant_forces(orc_blob *b, int n, const int *node, const double *force) {
    free(b->ctf_node);
    free(b->ctf_force);
    b->n_ctf = n;
    b->ctf_node = xcalloc(n > 0 ? n : 1, sizeof(int));
    b->ctf_force = xcalloc(3 * (size_t)(n > 0 ? n : 1), sizeof(double));
    if (n > 0) {
        memcpy(b->ctf_node, node, sizeof(int) * n);
        memcpy(b->ctf_force, force, sizeof(double) * 3 * n);
    }
}

/* Rotational constant forces (CTF 2) as Blob::load_ctforces stores them
 * (src/Blob.cpp:3246-3346), point + direction type only: axis [6n] = the axis point in meso
 * length units and the normalised direction; torque[i] 1 for 't', 0 for 'f'.  n = 0 clears. */
void orc_blob_set_rotational_forces(orc_blob *b, int n, const int *node, const double *F, const double *axis,
                                    const int *torque) {
    free(b->rctf_node);
    free(b->rctf_F);
    free(b->rctf_axis);
    free(b->rctf_torque);
    b->n_rctf = n;
    b->rctf_node = xcalloc(n > 0 ? n : 1, sizeof(int));
    b->rctf_F = xcalloc(n > 0 ? n : 1, sizeof(double));
    b->rctf_axis = xcalloc(6 * (size_t)(n > 0 ? n : 1), sizeof(double));
    b->rctf_torque = xcalloc(n > 0 ? n : 1, sizeof(int));
    if (n > 0) {
        memcpy(b->rctf_node, node, sizeof(int) * n);
        memcpy(b->rctf_F, F, sizeof(double) * n);
        memcpy(b->rctf_axis, axis, sizeof(double) * 6 * n);
        memcpy(b->rctf_torque, torque, sizeof(int) * n);
    }
}

/* arr3arr3VectorProduct (src/mat_vec_fns_II.cpp:127-134). */
static void cross_ref(const double *u, const double *v, double *w) {
    w[0] = u[1] * v[2] - v[1] * u[2];
    w[1] = -u[0] * v[2] + v[0] * u[2];
    w[2] = u[0] * v[1] - v[0] * u[1];
}

/* mag (src/mat_vec_fns_II.cpp:229-239). */
static double mag_ref(const double *v) {
    double s = 0.0;
    for (int i = 0; i < 3; i++) s += v[i] * v[i];
    return sqrt(s);
}

/* The force of one rotational constant force on its node at p (src/Blob.cpp:2104-2133): r from
 * p to its foot on the axis (intersectingPointToLine, src/mat_vec_fns_II.cpp:632-680, with
 * detByRows :261-269; a point within 3 eps of the axis is its own foot), f = r x d, scaled by F
 * ('f') or by F / |r| ('t'). */
static void rot_force(const double *p, const double *axis, double F, int torque, double *f) {
    const double *p1 = axis, *d = axis + 3;
    double p0p1[3], v2[3], cp[3], v1v2[3], r[3];
    for (int k = 0; k < 3; k++) p0p1[k] = p[k] - p1[k];
    cross_ref(p0p1, d, v2);
    if (mag_ref(v2) < 3.0 * DBL_EPSILON) {
        for (int k = 0; k < 3; k++) r[k] = p[k];
    } else {
        cross_ref(d, v2, cp);
        cross_ref(d, cp, v1v2);
        double c = 0;
        c = p0p1[0] * (d[1] * v1v2[2] - d[2] * v1v2[1]);
        c += p0p1[1] * (d[2] * v1v2[0] - d[0] * v1v2[2]);
        c += p0p1[2] * (d[0] * v1v2[1] - d[1] * v1v2[0]);
        c /= (v1v2[0] * v1v2[0] + v1v2[1] * v1v2[1] + v1v2[2] * v1v2[2]);
        for (int k = 0; k < 3; k++) cp[k] = c * cp[k];
        for (int k = 0; k < 3; k++) r[k] = p[k] + cp[k];
    }
    for (int k = 0; k < 3; k++) r[k] = r[k] - p[k];
    cross_ref(r, d, f);
    const double s = torque ? F / mag_ref(r) : F;
    for (int k = 0; k < 3; k++) f[k] = s * f[k];
}

/* rot_force for the tests (the checker of the checker compares it with the reference's own
 * helpers). */
void orc_rot_force(const double *p, const double *axis, double F, int torque, double *f) {
    rot_force(p, axis, F, torque, f);
}

/* Linear surface forces (src/Blob.cpp:3350-3424): the blob's surface faces [3 n_faces] (node
 * ids of Face::n[0..2]) and n_sets sets of face indices with their total force [3 n_sets]. */
void orc_blob_set_surface_forces(orc_blob *b, int n_faces, const int *faces, int n_sets, const int *set_off,
                                 const int *set_faces, const double *set_force) {
    free(b->face_nodes);
    free(b->set_off);
    free(b->set_faces);
    free(b->set_force);
    free(b->face_force);
    b->n_faces = n_faces;
    b->n_sets = n_sets;
    b->face_nodes = xcalloc(3 * (size_t)(n_faces > 0 ? n_faces : 1), sizeof(int));
    b->set_off = xcalloc((size_t)n_sets + 1, sizeof(int));
    b->set_faces = xcalloc(n_sets > 0 && set_off[n_sets] > 0 ? set_off[n_sets] : 1, sizeof(int));
    b->set_force = xcalloc(3 * (size_t)(n_sets > 0 ? n_sets : 1), sizeof(double));
    b->face_force = xcalloc(12 * (size_t)(n_faces > 0 ? n_faces : 1), sizeof(double));
    free(b->ssint_type);
    free(b->ssint_opp);
    free(b->face_geom);
    b->ssint_type = NULL;
    b->ssint_opp = NULL;
    b->face_geom = xcalloc(7 * (size_t)(n_faces > 0 ? n_faces : 1), sizeof(double));
    if (n_faces > 0) memcpy(b->face_nodes, faces, sizeof(int) * 3 * n_faces);
    if (n_sets > 0) {
        memcpy(b->set_off, set_off, sizeof(int) * (n_sets + 1));
        memcpy(b->set_faces, set_faces, sizeof(int) * set_off[n_sets]);
        memcpy(b->set_force, set_force, sizeof(double) * 3 * n_sets);
    }
}

/* Face::get_area (src/Face.cpp:244-261). */
static double face_area(const double *pos, const int *n) {
    double a[3], c[3], t[3];
    for (int k = 0; k < 3; k++) {
        a[k] = pos[3 * n[1] + k] - pos[3 * n[0] + k];
        c[k] = pos[3 * n[2] + k] - pos[3 * n[0] + k];
    }
    t[0] = a[1] * c[2] - a[2] * c[1];
    t[1] = a[2] * c[0] - a[0] * c[2];
    t[2] = a[0] * c[1] - a[1] * c[0];
    return .5 * sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
}

/* Blob::apply_ctforces, CTF 1 and CTF 3 (src/Blob.cpp:2094-2162): the linear forces go onto
 * the blob's force vector, the surface tractions onto the faces' node forces; rotational
 * forces (CTF 2) are not restated. */
static void blob_apply_ctforces(orc_blob *b) {
    for (int i = 0; i < b->n_ctf; i++)
        for (int k = 0; k < 3; k++) b->force[3 * b->ctf_node[i] + k] += b->ctf_force[3 * i + k];
    for (int i = 0; i < b->n_rctf; i++) {
        double f[3];
        const int n = b->rctf_node[i];
        rot_force(&b->pos[3 * n], &b->rctf_axis[6 * i], b->rctf_F[i], b->rctf_torque[i], f);
        for (int k = 0; k < 3; k++) b->force[3 * n + k] += f[k];
    }
    for (int s = 0; s < b->n_sets; s++) {
        int j0 = b->set_off[s], j1 = b->set_off[s + 1];
        double *areas = xcalloc(j1 > j0 ? j1 - j0 : 1, sizeof(double));
        double total = 0;
        for (int j = j0; j < j1; j++) {
            areas[j - j0] = face_area(b->pos, &b->face_nodes[3 * b->set_faces[j]]);
            total += areas[j - j0];
        }
        for (int j = j0; j < j1; j++) {
            double t[3], f = areas[j - j0] / (3 * total);
            for (int k = 0; k < 3; k++) t[k] = f * b->set_force[3 * s + k];
            double *ff = &b->face_force[12 * b->set_faces[j]];
            for (int v = 0; v < 3; v++)
                for (int k = 0; k < 3; k++) ff[3 * v + k] += t[k];
        }
        free(areas);
    }
}

/* Springs of World::apply_springs (src/World.cpp:3808-3846) over the blobs of one World:
 * spring i joins node node0[i] of blob blob0[i] to node node1[i] of blob blob1[i]; k and l in
 * meso units (World::load_springs, src/World.cpp:3509-3510).  n = 0 clears. */
static int g_nspr;
static int *g_spr_b0, *g_spr_n0, *g_spr_b1, *g_spr_n1;
static double *g_spr_k, *g_spr_l;
void orc_world_set_springs(int n, const int *blob0, const int *node0, const int *blob1, const int *node1,
                           const double *k, const double *l) {
    free(g_spr_b0);
    free(g_spr_n0);
    free(g_spr_b1);
    free(g_spr_n1);
    free(g_spr_k);
    free(g_spr_l);
    g_nspr = n;
    size_t m = n > 0 ? n : 1;
    g_spr_b0 = xcalloc(m, sizeof(int));
    g_spr_n0 = xcalloc(m, sizeof(int));
    g_spr_b1 = xcalloc(m, sizeof(int));
    g_spr_n1 = xcalloc(m, sizeof(int));
    g_spr_k = xcalloc(m, sizeof(double));
    g_spr_l = xcalloc(m, sizeof(double));
    if (n > 0) {
        memcpy(g_spr_b0, blob0, sizeof(int) * n);
        memcpy(g_spr_n0, node0, sizeof(int) * n);
        memcpy(g_spr_b1, blob1, sizeof(int) * n);
        memcpy(g_spr_n1, node1, sizeof(int) * n);
        memcpy(g_spr_k, k, sizeof(double) * n);
        memcpy(g_spr_l, l, sizeof(double) * n);
    }
}

/* mag (src/mat_vec_fns_II.cpp:229-239). */
static double mag3(const double v[3]) {
    double s = 0.0;
    for (int i = 0; i < 3; i++) s += v[i] * v[i];
    return sqrt(s);
}

static void world_apply_springs(orc_blob **blobs) {
    for (int i = 0; i < g_nspr; i++) {
        const double *p0 = &blobs[g_spr_b0[i]]->pos[3 * g_spr_n0[i]], *p1 = &blobs[g_spr_b1[i]]->pos[3 * g_spr_n1[i]];
        double sep[3], sn[3], norm = 0.0;
        for (int k = 0; k < 3; k++) sep[k] = p1[k] - p0[k];
        /* arr3Normalise2 (src/mat_vec_fns_II.cpp:164-176): a zero separation gives a zero direction */
        for (int k = 0; k < 3; k++) norm += sep[k] * sep[k];
        if (norm == 0.0) {
            sn[0] = sn[1] = sn[2] = 0.0;
        } else {
            norm = sqrt(norm);
            for (int k = 0; k < 3; k++) sn[k] = sep[k] / norm;
        }
        double fm = g_spr_k[i] * (mag3(sep) - g_spr_l[i]);
        double *f0 = &blobs[g_spr_b0[i]]->force[3 * g_spr_n0[i]], *f1 = &blobs[g_spr_b1[i]]->force[3 * g_spr_n1[i]];
        for (int k = 0; k < 3; k++) f0[k] += fm * sn[k];
        for (int k = 0; k < 3; k++) f1[k] += -1 * fm * sn[k];
    }
}

/* Spring energy of World::make_measurements (src/World.cpp:4316-4329) and
 * get_spring_field_energy(-1, -1) (:3848-3858): summed per (blob0, blob1) pair in spring
 * order, then over the pairs. */
double orc_world_spring_energy(orc_blob **blobs, int nblobs) {
    double *e = xcalloc((size_t)nblobs * nblobs, sizeof(double));
    for (int i = 0; i < g_nspr; i++) {
        const double *a = &blobs[g_spr_b0[i]]->pos[3 * g_spr_n0[i]], *c = &blobs[g_spr_b1[i]]->pos[3 * g_spr_n1[i]];
        double d[3];
        for (int k = 0; k < 3; k++) d[k] = a[k] - c[k];
        double m = mag3(d);
        e[g_spr_b0[i] * nblobs + g_spr_b1[i]] += 0.5 * g_spr_k[i] * (m - g_spr_l[i]) * (m - g_spr_l[i]);
    }
    double tot = 0.0;
    for (int i = 0; i < nblobs * nblobs; i++) tot += e[i];
    free(e);
    return tot;
}

/* Blob::update_internal_forces (src/Blob.cpp:638-735), single thread. Returns #inversions. */
int orc_blob_update_internal_forces(orc_blob *b, orc_stream *g) {
    int inversions = 0;
    blob_apply_ctforces(b); /* src/Blob.cpp:647 */
    for (int n = 0; n < b->num_elements; n++) {
        m3 J, stress, F;
        double du[12];
        const int *t = &b->topo[10 * n];
        double *dpsi = &b->dpsi[12 * n];
        elem_jacobian(b->pos, t, J);
        if (elem_dpsi_vol(J, dpsi, &b->vol[n], &b->last_det[n]) != 0) inversions++;
        elem_viscosity(dpsi, b->A[n], b->B[n], b->vol[n], (double(*)[12]) & b->visc[144 * n]);
        memset(stress, 0, sizeof(stress));
        elem_shear_stress(J, &b->J_inv_0[9 * n], b->G[n], b->vol_0[n], b->vol[n], F, stress);
        memcpy(&b->F_ij[9 * n], F, sizeof(F));
        elem_bulk_stress(b->G[n], b->E[n], b->vol_0[n], b->vol[n], stress);
        if (b->calc_noise == 1) elem_fluct_stress(b->kT, b->dt, b->A[n], b->B[n], b->vol[n], g, stress);
        /* mat3_double_contraction_symmetric (src/mat_vec_fns.cpp:131-134) */
        b->stress_mag[n] = sqrt(stress[0][0] * stress[0][0] + stress[1][1] * stress[1][1] +
                                stress[2][2] * stress[2][2] +
                                2 * (stress[0][1] * stress[0][1] + stress[0][2] * stress[0][2] +
                                     stress[1][2] * stress[1][2]));
        if (b->linear_solver != 3) {
            /* mass-ful solvers: viscous force K_e u_e (src/Blob.cpp:703-706) from
             * get_element_velocity_vector (:454-469), x-major over the 4 vertices, and
             * mat12_apply (src/mat_vec_fns.cpp:26-34) */
            double v[12];
            const double(*V)[12] = (const double(*)[12]) & b->visc[144 * n];
            for (int i = 0; i < 3; i++)
                for (int a = 0; a < 4; a++) v[4 * i + a] = b->vel[3 * t[a] + i];
            for (int i = 0; i < 12; i++) {
                double acc = 0;
                for (int j = 0; j < 12; j++) acc += V[i][j] * v[j];
                du[i] = acc;
            }
        } else {
            for (int k = 0; k < 12; k++) du[k] = 0;
        }
        /* apply_stress_tensor (:441-449) */
        double vol = b->vol[n];
        for (int i = 0; i < 3; i++)
            for (int a = 0; a < 4; a++)
                du[4 * i + a] +=
                    vol * (dpsi[a] * stress[i][0] + dpsi[4 + a] * stress[i][1] + dpsi[8 + a] * stress[i][2]);
        /* add_element_force_vector (:474-490): node_force[a][i] -= du[4i+a] */
        double *nf = &b->node_force[30 * n];
        for (int i = 0; i < 3; i++)
            for (int a = 0; a < 4; a++) nf[3 * a + i] -= du[4 * i + a];
    }
    b->last_inversions = inversions;
    return inversions;
}

/* SparseMatrixFixedPattern::build + calc_inverse_diagonal. */
static void blob_build_matrix(orc_blob *b) {
    for (int k = 0; k < b->nnz; k++) {
        double sum = 0.0;
        for (int s = b->src_off[k]; s < b->src_off[k + 1]; s++) sum += *(b->src[s]);
        b->val[k] = sum;
    }
    for (int i = 0; i < b->num_rows; i++) b->diag_inv[i] = 1.0 / b->val[b->diag_idx[i]];
}

/* SparseMatrixFixedPattern::apply(vector3*, vector3*) (src/SparseMatrixFixedPattern.cpp:148-191). */
static void blob_spmv(const orc_blob *b, const double *in, double *out) {
    for (int i = 0; i < b->num_rows; i++) {
        double w = 0.0;
        for (int j = b->key[i]; j < b->key[i + 1]; ++j) w += b->val[j] * in[b->col[j]];
        out[i] = w;
    }
}

/* NoMassCGSolver::solve (src/NoMassCGSolver.cpp:165-197) with its vector kernels (:205-292)
 * and vec3_add_to_scaled / vec3_scale_and_add (src/mat_vec_fns.cpp:250-275). */
static int blob_solve(orc_blob *b, double *x) {
    int N = b->num_nodes;
    double *r = b->r, *p = b->p, *z = b->z, *q = b->q, *f = b->f, *D = b->diag_inv;
    blob_build_matrix(b);
    double delta_new = 0, delta_old, pTq, alpha;
    for (int i = 0; i < N; i++) {
        for (int c = 0; c < 3; c++) {
            r[3 * i + c] = x[3 * i + c];
            f[3 * i + c] = x[3 * i + c];
            x[3 * i + c] = 0;
            z[3 * i + c] = D[3 * i + c] * r[3 * i + c];
            p[3 * i + c] = z[3 * i + c];
        }
        delta_new += r[3 * i] * z[3 * i] + r[3 * i + 1] * z[3 * i + 1] + r[3 * i + 2] * z[3 * i + 2];
    }
    for (int it = 0; it < b->max_iterations_cg; it++) {
        blob_spmv(b, p, q);
        pTq = 0;
        for (int i = 0; i < N; ++i) pTq += p[3 * i] * q[3 * i] + p[3 * i + 1] * q[3 * i + 1] + p[3 * i + 2] * q[3 * i + 2];
        alpha = delta_new / pTq;
        for (int i = 0; i < 3 * N; i++) x[i] += alpha * p[i];
        for (int i = 0; i < 3 * N; i++) r[i] += -alpha * q[i];
        double r2 = 0, f2 = 0, res;
        for (int i = 0; i < N; i++) {
            r2 += r[3 * i] * r[3 * i] + r[3 * i + 1] * r[3 * i + 1] + r[3 * i + 2] * r[3 * i + 2];
            f2 += f[3 * i] * f[3 * i] + f[3 * i + 1] * f[3 * i + 1] + f[3 * i + 2] * f[3 * i + 2];
        }
        res = (f2 == 0.0) ? 0.0 : r2 / f2;
        if (res < b->epsilon2) {
            b->last_cg_iterations = it + 1;
            b->last_cg_status = 0;
            return 0;
        }
        delta_old = delta_new;
        delta_new = 0;
        for (int i = 0; i < N; i++) {
            z[3 * i] = D[3 * i] * r[3 * i];
            z[3 * i + 1] = D[3 * i + 1] * r[3 * i + 1];
            z[3 * i + 2] = D[3 * i + 2] * r[3 * i + 2];
            delta_new += r[3 * i] * z[3 * i] + r[3 * i + 1] * z[3 * i + 1] + r[3 * i + 2] * z[3 * i + 2];
        }
        double beta = delta_new / delta_old;
        for (int i = 0; i < 3 * N; i++) p[i] = beta * p[i] + z[i];
    }
    b->last_cg_iterations = b->max_iterations_cg;
    b->last_cg_status = -1;
    return -1;
}

/* Blob::aggregate_forces_and_solve, nomass branch (src/Blob.cpp:3826-3932).  Leaves the
 * RHS in b->f (copied by the solver) and the velocity solution in b->force. */
static int blob_aggregate_forces(orc_blob *b, orc_stream *g) {
    int N = b->num_nodes;
    for (int n = 0; n < N; n++)
        for (int m = b->contrib_off[n]; m < b->contrib_off[n + 1]; m++) {
            const double *c = &b->node_force[3 * b->contrib[m]];
            b->force[3 * n] += c[0];
            b->force[3 * n + 1] += c[1];
            b->force[3 * n + 2] += c[2];
        }
    /* surface face forces (src/Blob.cpp:3842-3851), face by face, Face::n[0..3]; the opposite
     * vertex n[3] (known once interaction types are set) only carries steric forces */
    for (int n = 0; n < b->n_faces; n++) {
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) b->force[3 * b->face_nodes[3 * n + i] + k] += b->face_force[12 * n + 3 * i + k];
        if (b->ssint_opp && b->ssint_opp[n] >= 0)
            for (int k = 0; k < 3; k++) b->force[3 * b->ssint_opp[n] + k] += b->face_force[12 * n + 9 + k];
    }
    if (b->calc_stokes == 1 && b->linear_solver != 3) {
        /* mass-ful branch (src/Blob.cpp:3854-3877): drag on the node velocity, then noise */
        for (int i = 0; i < N; i++) {
            for (int c = 0; c < 3; c++) b->force[3 * i + c] -= b->vel[3 * i + c] * b->stokes_drag[i];
            if (b->calc_noise == 1)
                for (int c = 0; c < 3; c++)
                    b->force[3 * i + c] -= rand_half(g) * sqrt((24 * b->kT * b->stokes_drag[i]) / (b->dt));
        }
    } else if (b->calc_stokes == 1 && b->calc_noise == 1)
        for (int i = 0; i < N; i++)
            for (int c = 0; c < 3; c++)
                b->force[3 * i + c] -= rand_half(g) * sqrt((24 * b->kT * b->stokes_drag[i]) / (b->dt));
    /* Blob::linearise_force (src/Blob.cpp:1066-1085) */
    for (int e = 0; e < b->num_elements; e++) {
        const int *nI = &b->topo[10 * e];
        double *F = b->force;
        for (int j = 0; j < 3; j++) {
            F[3 * nI[0] + j] += 0.5 * (F[3 * nI[4] + j] + F[3 * nI[5] + j] + F[3 * nI[6] + j]);
            F[3 * nI[1] + j] += 0.5 * (F[3 * nI[4] + j] + F[3 * nI[7] + j] + F[3 * nI[8] + j]);
            F[3 * nI[2] + j] += 0.5 * (F[3 * nI[5] + j] + F[3 * nI[7] + j] + F[3 * nI[9] + j]);
            F[3 * nI[3] + j] += 0.5 * (F[3 * nI[6] + j] + F[3 * nI[8] + j] + F[3 * nI[9] + j]);
        }
        for (int j = 4; j < 10; j++) F[3 * nI[j]] = F[3 * nI[j] + 1] = F[3 * nI[j] + 2] = 0.;
    }
    for (int n = 0; n < b->num_pinned; n++) {
        int pn = b->pinned[n];
        b->force[3 * pn] = b->force[3 * pn + 1] = b->force[3 * pn + 2] = 0;
    }
    return 0;
}

/* ---- mass-ful solvers ------------------------------------------------------ */

/* ConjugateGradientSolver::init (src/ConjugateGradientSolver.cpp:62-203): the consistent
 * mass matrix of the linear vertices (0.1 rho vol_0 on the diagonal, 0.05 off it, summed in
 * element order), 1 on the diagonal of pinned rows and of midpoints, the nonzeros of each
 * dense row in ascending column order, and the inverse diagonal as preconditioner.  The
 * dense N_t x N_t scan of the reference is restated row by row over a sorted column list. */
static void mass_cg_init(orc_blob *b) {
    int N = b->num_nodes, E = b->num_elements;
    /* per-row (col, value) lists, built in the reference's accumulation order */
    int *cnt = xcalloc(N, sizeof(int));
    for (int e = 0; e < E; e++)
        for (int i = 0; i < 10; i++) cnt[b->topo[10 * e + i]] += 10;
    int *off = xcalloc((size_t)N + 1, sizeof(int));
    for (int i = 0; i < N; i++) off[i + 1] = off[i] + cnt[i];
    int *cols = xcalloc((size_t)off[N], sizeof(int));
    double *vals = xcalloc((size_t)off[N], sizeof(double));
    int *len = xcalloc(N, sizeof(int));
    for (int e = 0; e < E; e++)
        for (int i = 0; i < 10; i++)
            for (int j = 0; j < 10; j++) {
                int ni = b->topo[10 * e + i], nj = b->topo[10 * e + j], k;
                if (!(i < 4 && j < 4) && i != j) continue;
                for (k = 0; k < len[ni]; k++)
                    if (cols[off[ni] + k] == nj) break;
                if (k == len[ni]) {
                    cols[off[ni] + k] = nj;
                    vals[off[ni] + k] = 0;
                    len[ni]++;
                }
                double *m = &vals[off[ni] + k];
                if (i < 4 && j < 4) {
                    if (b->is_pinned[ni] == 0 && b->is_pinned[nj] == 0)
                        *m += (i == j ? .1 : .05) * b->rho[e] * b->vol_0[e];
                    else if (i == j)
                        *m = 1;
                } else {
                    *m = 1; /* i == j, a midpoint */
                }
            }
    /* nonzeros in ascending column order */
    int nnz = 0;
    for (int i = 0; i < N; i++) {
        int *c = &cols[off[i]];
        double *v = &vals[off[i]];
        for (int a = 1; a < len[i]; a++)
            for (int q = a; q > 0 && c[q - 1] > c[q]; q--) {
                int tc = c[q];
                c[q] = c[q - 1];
                c[q - 1] = tc;
                double tv = v[q];
                v[q] = v[q - 1];
                v[q - 1] = tv;
            }
        for (int a = 0; a < len[i]; a++) nnz += v[a] != 0;
    }
    b->mkey = xcalloc((size_t)N + 1, sizeof(int));
    b->mcol = xcalloc(nnz, sizeof(int));
    b->mval = xcalloc(nnz, sizeof(double));
    b->mprec = xcalloc(N, sizeof(double));
    int k = 0;
    for (int i = 0; i < N; i++) {
        b->mkey[i] = k;
        double diag = 0;
        for (int a = 0; a < len[i]; a++) {
            double v = vals[off[i] + a];
            if (cols[off[i] + a] == i) diag = v;
            if (v == 0) continue;
            b->mcol[k] = cols[off[i] + a];
            b->mval[k] = v;
            k++;
        }
        b->mprec[i] = 1.0 / diag;
    }
    b->mkey[N] = k;
    free(cnt);
    free(off);
    free(cols);
    free(vals);
    free(len);
}

/* MassLumpedSolver::init (src/MassLumpedSolver.cpp:40-79). */
static void mass_lumped_init(orc_blob *b) {
    int N = b->num_nodes;
    b->inv_M = xcalloc(N, sizeof(double));
    for (int e = 0; e < b->num_elements; e++)
        for (int i = 0; i < 10; i++) {
            int ni = b->topo[10 * e + i];
            if (i < 4)
                b->inv_M[ni] += .25 * b->rho[e] * b->vol_0[e];
            else
                b->inv_M[ni] = 1;
        }
    for (int i = 0; i < b->num_pinned; i++) b->inv_M[b->pinned[i]] = 1.0;
    for (int i = 0; i < N; i++) b->inv_M[i] = 1.0 / b->inv_M[i];
}

/* Selects the Blob's linear solver (Blob::init, src/Blob.cpp:358-373): 1 ConjugateGradient
 * (mass matrix), 2 MassLumped, 3 NoMassCG.  Returns 0, or -1 for another value. */
int orc_blob_set_solver(orc_blob *b, int solver) {
    if (solver < 1 || solver > 3) return -1;
    b->linear_solver = solver;
    if (solver == 1 && !b->mkey) mass_cg_init(b);
    if (solver == 2 && !b->inv_M) mass_lumped_init(b);
    return 0;
}

/* ConjugateGradientSolver::solve (src/ConjugateGradientSolver.cpp:205-230) with
 * conjugate_gradient_residual_assume_x_zero (:241-265), parallel_sparse_matrix_apply
 * (:268-287), parallel_vector_add_self / parallel_vector_add (:289-311),
 * parallel_apply_preconditioner (:314-329) and residual2 (:332-349): every sum in row order. */
static int mass_cg_solve(orc_blob *b, double *x) {
    int N = b->num_nodes;
    double *r = b->r, *d = b->p, *s = b->z, *q = b->q, *f = b->f, *P = b->mprec;
    double delta_new = 0, delta_old, dTq, alpha;
    for (int i = 0; i < N; i++) {
        for (int c = 0; c < 3; c++) {
            r[3 * i + c] = x[3 * i + c];
            f[3 * i + c] = x[3 * i + c];
            x[3 * i + c] = 0;
            d[3 * i + c] = P[i] * r[3 * i + c];
        }
        delta_new += r[3 * i] * d[3 * i] + r[3 * i + 1] * d[3 * i + 1] + r[3 * i + 2] * d[3 * i + 2];
    }
    for (int it = 0; it < b->max_iterations_cg; it++) {
        double r2 = 0, f2 = 0;
        for (int i = 0; i < N; i++) {
            r2 += r[3 * i] * r[3 * i] + r[3 * i + 1] * r[3 * i + 1] + r[3 * i + 2] * r[3 * i + 2];
            f2 += f[3 * i] * f[3 * i] + f[3 * i + 1] * f[3 * i + 1] + f[3 * i + 2] * f[3 * i + 2];
        }
        double res = f2 == 0.0 ? 0.0 : r2 / f2;
        if (res < b->epsilon2) {
            b->last_cg_iterations = it;
            b->last_cg_status = 0;
            return 0;
        }
        dTq = 0;
        for (int i = 0; i < N; i++) {
            double qx = 0, qy = 0, qz = 0;
            for (int j = b->mkey[i]; j < b->mkey[i + 1]; j++) {
                int c = b->mcol[j];
                qx += b->mval[j] * d[3 * c];
                qy += b->mval[j] * d[3 * c + 1];
                qz += b->mval[j] * d[3 * c + 2];
            }
            q[3 * i] = qx;
            q[3 * i + 1] = qy;
            q[3 * i + 2] = qz;
            dTq += d[3 * i] * qx + d[3 * i + 1] * qy + d[3 * i + 2] * qz;
        }
        alpha = delta_new / dTq;
        for (int i = 0; i < 3 * N; i++) x[i] += alpha * d[i];
        for (int i = 0; i < 3 * N; i++) r[i] += -alpha * q[i];
        delta_old = delta_new;
        delta_new = 0;
        for (int i = 0; i < N; i++) {
            for (int c = 0; c < 3; c++) s[3 * i + c] = P[i] * r[3 * i + c];
            delta_new += r[3 * i] * s[3 * i] + r[3 * i + 1] * s[3 * i + 1] + r[3 * i + 2] * s[3 * i + 2];
        }
        for (int i = 0; i < 3 * N; i++) d[i] = s[i] + (delta_new / delta_old) * d[i];
    }
    b->last_cg_iterations = b->max_iterations_cg;
    b->last_cg_status = -1;
    return -1;
}

/* MassLumpedSolver::solve (src/MassLumpedSolver.cpp:82-90). */
static int mass_lumped_solve(orc_blob *b, double *x) {
    memcpy(b->f, x, sizeof(double) * 3 * (size_t)b->num_nodes); /* kept for comparisons only */
    for (int i = 0; i < b->num_nodes; i++)
        for (int c = 0; c < 3; c++) x[3 * i + c] = x[3 * i + c] * b->inv_M[i];
    b->last_cg_iterations = 0;
    b->last_cg_status = 0;
    return 0;
}

static int blob_solver_solve(orc_blob *b, double *x) {
    if (b->linear_solver == 1) return mass_cg_solve(b, x);
    if (b->linear_solver == 2) return mass_lumped_solve(b, x);
    return blob_solve(b, x);
}

/* Overwrite all N_t node velocities (a mass-ful restart frame). */
void orc_blob_set_velocities(orc_blob *b, const double *vel) {
    memcpy(b->vel, vel, sizeof(double) * 3 * (size_t)b->num_nodes);
}

/* Kinetic energy of Blob::make_measurements (src/Blob.cpp:1385-1407, :1436): per element
 * u_e . (M_e u_e) with apply_element_mass_matrix (src/tetra_element_linear.cpp:534-553),
 * halved.  Zero for the massless solver. */
double orc_blob_kinetic_energy(const orc_blob *b) {
    static const double M[4][4] = {{.1, .05, .05, .05}, {.05, .1, .05, .05}, {.05, .05, .1, .05}, {.05, .05, .05, .1}};
    if (b->linear_solver == 3) return 0.0;
    double ke = 0.0;
    for (int n = 0; n < b->num_elements; n++) {
        const int *t = &b->topo[10 * n];
        double u[12], mu[12] = {0};
        for (int i = 0; i < 3; i++)
            for (int a = 0; a < 4; a++) u[4 * i + a] = b->vel[3 * t[a] + i];
        for (int k = 0; k < 12; k += 4)
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) mu[i + k] += M[i][j] * u[j + k];
        for (int i = 0; i < 12; i++) mu[i] = mu[i] * (b->vol_0[n] * b->rho[n]);
        ke += u[0] * mu[0] + u[1] * mu[1] + u[2] * mu[2] + u[3] * mu[3] + u[4] * mu[4] + u[5] * mu[5] +
              u[6] * mu[6] + u[7] * mu[7] + u[8] * mu[8] + u[9] * mu[9] + u[10] * mu[10] + u[11] * mu[11];
    }
    return ke * 0.5;
}

/* Blob::update_positions (src/Blob.cpp:737-764) with euler_integrate (:3937-3967). */
int orc_blob_update_positions(orc_blob *b, orc_stream *g) {
    blob_aggregate_forces(b, g);
    int st = blob_solver_solve(b, b->force);
    if (st != 0) return st; /* the reference skips integration on a solver error */
    if (b->linear_solver == 3) {
        for (int i = 0; i < b->num_nodes; i++)
            for (int c = 0; c < 3; c++) {
                b->vel[3 * i + c] = b->force[3 * i + c];
                b->pos[3 * i + c] += b->force[3 * i + c] * b->dt;
            }
    } else {
        for (int i = 0; i < b->num_nodes; i++)
            for (int c = 0; c < 3; c++) {
                b->vel[3 * i + c] += b->force[3 * i + c] * b->dt;
                b->pos[3 * i + c] += b->vel[3 * i + c] * b->dt;
            }
    }
    for (int e = 0; e < b->num_elements; e++) elem_linearise(b->pos, &b->topo[10 * e]);
    return 0;
}

/* One World::run iteration over nblobs blobs sharing one stream (src/World.cpp:1844-2043):
 * zero forces, element phase for every blob, then position phase for every blob.
 * Returns the total inversion count (the reference aborts the run if > 0), or -1 if a
 * solve failed. */
/* World::run's periodic box (src/World.cpp:1850-1901): box_dim and which walls are periodic,
 * bits 0..5 = wall_x_1, wall_x_2, wall_y_1, wall_y_2, wall_z_1, wall_z_2.  mask 0: no walls
 * act (the default here, for the tests that drive blobs without a World box). */
static double g_box[3];
static int g_box_mask;
void orc_world_set_box(const double *box, int mask) {
    for (int k = 0; k < 3; k++) g_box[k] = box[k];
    g_box_mask = mask;
}

/* Blob::get_centroid (src/Blob.cpp:971-983) over all N_t nodes. */
void orc_blob_centroid(const orc_blob *b, double *com) {
    com[0] = com[1] = com[2] = 0;
    for (int n = 0; n < b->num_nodes; n++)
        for (int k = 0; k < 3; k++) com[k] += b->pos[3 * n + k];
    for (int k = 0; k < 3; k++) com[k] /= b->num_nodes;
}

/* A blob whose centroid has left the box through a periodic wall moves by one box length
 * back in (Blob::move, src/Blob.cpp:935-949: every node); per axis the low wall is checked
 * first, the high one only if the centroid is not below zero. */
static void blob_wrap(orc_blob *b) {
    if (!g_box_mask) return;
    double com[3], d[3] = {0, 0, 0};
    int move = 0;
    orc_blob_centroid(b, com);
    for (int k = 0; k < 3; k++) {
        if (com[k] < 0) {
            if (g_box_mask & (1 << (2 * k))) {
                d[k] += g_box[k];
                move = 1;
            }
        } else if (com[k] > g_box[k]) {
            if (g_box_mask & (2 << (2 * k))) {
                d[k] -= g_box[k];
                move = 1;
            }
        }
    }
    if (move)
        for (int n = 0; n < b->num_nodes; n++)
            for (int k = 0; k < 3; k++) b->pos[3 * n + k] += d[k];
}

/* ---- surface-surface interactions: the Lennard-Jones VdW_solver ---------------------- */

/* VdW_solver's three-point triangle rule (src/VdW_solver.cpp:62-73): weight, then the
 * barycentric coordinates, as the decimal constants the reference writes. */
static const double g_gw = 0.333333333333333;
static const double g_geta[3][3] = {{0.666666666666667, 0.166666666666667, 0.166666666666667},
                                    {0.166666666666667, 0.666666666666667, 0.166666666666667},
                                    {0.166666666666667, 0.166666666666667, 0.666666666666667}};
/* adjacent_cell_lookup_table (src/VdW_solver.cpp:31-59): dx, dy, dz each -1..1, z fastest */
static const int g_adj[27][3] = {{-1, -1, -1}, {-1, -1, 0}, {-1, -1, 1}, {-1, 0, -1}, {-1, 0, 0}, {-1, 0, 1},
                                 {-1, 1, -1},  {-1, 1, 0},  {-1, 1, 1},  {0, -1, -1}, {0, -1, 0}, {0, -1, 1},
                                 {0, 0, -1},   {0, 0, 0},   {0, 0, 1},   {0, 1, -1},  {0, 1, 0},  {0, 1, 1},
                                 {1, -1, -1},  {1, -1, 0},  {1, -1, 1},  {1, 0, -1},  {1, 0, 0},  {1, 0, 1},
                                 {1, 1, -1},   {1, 1, 0},   {1, 1, 1}};

/* SSINT_matrix in meso units (LJI(i, j) = i * ntypes + j), the lookup grid and the solver's
 * switches (World::init, src/World.cpp:896-970; VdW_solver::init). ntypes 0: off.  kind: 1
 * lennard-jones, 2 ljsteric (its soft-core branch below Rmin), 3 gensoft, 4 steric (volume
 * exclusion only); 2-4 apply the steric forces where face elements overlap. */
static int g_ss_ntypes, g_ss_inc_self, g_ss_es_update, g_ss_es_count, g_ss_N[3], g_ss_kind = 1;
static double *g_ss_emin, *g_ss_rmin, *g_ss_k0, g_ss_h;
static double *g_ss_energy; /* fieldenergy [nblobs * nblobs] of the last solve */
static int g_ss_nblobs;
/* the lookup grid: per pool entry (blob, face) its cell and the next entry down its stack */
static int g_ss_npool, *g_ss_pool_blob, *g_ss_pool_face, *g_ss_pool_cell, *g_ss_next, *g_ss_root;

void orc_world_set_ssint(int kind, int ntypes, const double *emin, const double *rmin, const double *k0, const int *N,
                         double h, int inc_self, int es_update) {
    free(g_ss_emin);
    free(g_ss_rmin);
    free(g_ss_k0);
    free(g_ss_root);
    g_ss_emin = g_ss_rmin = g_ss_k0 = NULL;
    g_ss_root = NULL;
    g_ss_ntypes = ntypes;
    g_ss_kind = kind;
    if (ntypes <= 0) return;
    g_ss_emin = xcalloc((size_t)ntypes * ntypes, sizeof(double));
    g_ss_rmin = xcalloc((size_t)ntypes * ntypes, sizeof(double));
    g_ss_k0 = xcalloc((size_t)ntypes * ntypes, sizeof(double));
    memcpy(g_ss_emin, emin, sizeof(double) * ntypes * ntypes);
    memcpy(g_ss_rmin, rmin, sizeof(double) * ntypes * ntypes);
    memcpy(g_ss_k0, k0, sizeof(double) * ntypes * ntypes);
    for (int k = 0; k < 3; k++) g_ss_N[k] = N[k];
    g_ss_h = h;
    g_ss_inc_self = inc_self;
    g_ss_es_update = es_update;
    g_ss_es_count = es_update; /* World::run starts with an update (src/World.cpp:1805) */
    g_ss_root = xcalloc((size_t)N[0] * N[1] * N[2], sizeof(int));
}

/* Blob::load_ssint (src/Blob.cpp:3494-3578): the type of every surface face (-1 inactive);
 * call after orc_blob_set_surface_forces has given the faces. */
void orc_blob_set_ssint_types(orc_blob *b, const int *types, const int *opposite) {
    free(b->ssint_type);
    free(b->ssint_opp);
    b->ssint_type = xcalloc(b->n_faces > 0 ? b->n_faces : 1, sizeof(int));
    b->ssint_opp = xcalloc(b->n_faces > 0 ? b->n_faces : 1, sizeof(int));
    if (b->n_faces > 0) {
        memcpy(b->ssint_type, types, sizeof(int) * b->n_faces);
        memcpy(b->ssint_opp, opposite, sizeof(int) * b->n_faces);
    }
}

/* Face::calc_area_normal_centroid (src/Face.cpp:197-219). */
static void face_geometry(const double *pos, const int *fn, double *geo) {
    const double *p0 = &pos[3 * fn[0]], *p1 = &pos[3 * fn[1]], *p2 = &pos[3 * fn[2]];
    double a[3], c[3], nrm[3];
    for (int k = 0; k < 3; k++) {
        a[k] = p1[k] - p0[k];
        c[k] = p2[k] - p0[k];
    }
    nrm[0] = a[1] * c[2] - a[2] * c[1];
    nrm[1] = a[2] * c[0] - a[0] * c[2];
    nrm[2] = a[0] * c[1] - a[1] * c[0];
    const double m = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    geo[6] = .5 * m;
    for (int k = 0; k < 3; k++) geo[3 + k] = nrm[k] / m;
    for (int k = 0; k < 3; k++) geo[k] = (1.0 / 3.0) * (p0[k] + p1[k] + p2[k]);
}

/* LinkedListCube::pbc (src/LinkedListCube.tpp:282-299): one period at most. */
static void cell_pbc(int *x) {
    for (int k = 0; k < 3; k++) {
        if (x[k] < 0)
            x[k] += g_ss_N[k];
        else if (x[k] >= g_ss_N[k])
            x[k] -= g_ss_N[k];
    }
}

/* The es_update refresh (src/World.cpp:1821-1837, 1907-1925): every face's geometry, then
 * NearestNeighbourLinkedListCube::build_nearest_neighbour_lookup
 * (src/NearestNeighbourLinkedListCube.cpp:26-77): the pool is the active faces blob by blob in
 * face order (src/World.cpp:940-962); entry i goes on top of its cell's stack.  Returns -1 for a
 * centroid outside the grid after one period. */
static int ssint_refresh(orc_blob **blobs, int nblobs) {
    int n = 0;
    for (int b = 0; b < nblobs; b++)
        for (int f = 0; f < blobs[b]->n_faces; f++) {
            face_geometry(blobs[b]->pos, &blobs[b]->face_nodes[3 * f], &blobs[b]->face_geom[7 * f]);
            if (blobs[b]->ssint_type && blobs[b]->ssint_type[f] != -1) n++;
        }
    free(g_ss_pool_blob);
    free(g_ss_pool_face);
    free(g_ss_pool_cell);
    free(g_ss_next);
    g_ss_pool_blob = xcalloc(n ? n : 1, sizeof(int));
    g_ss_pool_face = xcalloc(n ? n : 1, sizeof(int));
    g_ss_pool_cell = xcalloc(n ? n : 1, sizeof(int));
    g_ss_next = xcalloc(n ? n : 1, sizeof(int));
    g_ss_npool = n;
    const int ncell = g_ss_N[0] * g_ss_N[1] * g_ss_N[2];
    for (int c = 0; c < ncell; c++) g_ss_root[c] = -1;
    int i = 0;
    for (int b = 0; b < nblobs; b++)
        for (int f = 0; f < blobs[b]->n_faces; f++) {
            if (!blobs[b]->ssint_type || blobs[b]->ssint_type[f] == -1) continue;
            const double *geo = &blobs[b]->face_geom[7 * f];
            int x[3];
            for (int k = 0; k < 3; k++) x[k] = (int)floor(geo[k] / g_ss_h);
            cell_pbc(x);
            for (int k = 0; k < 3; k++)
                if (x[k] < 0 || x[k] >= g_ss_N[k]) return -1;
            const int c = x[0] * g_ss_N[1] * g_ss_N[2] + x[1] * g_ss_N[2] + x[2];
            g_ss_pool_blob[i] = b;
            g_ss_pool_face[i] = f;
            g_ss_pool_cell[i] = c;
            g_ss_next[i] = g_ss_root[c];
            g_ss_root[c] = i;
            i++;
        }
    return 0;
}

/* VdW_solver::calc_ljinterpolated_factors (src/VdW_solver.cpp:774-784): the soft core below
 * Rmin that ljsteric uses. */
static void lj_soft_factors(double mag_r, double Emin, double Rmini, double *force_mag, double *e) {
    const double f = mag_r * Rmini, f2 = f * f;
    *e = g_gw * g_gw * Emin * f2 * (2 * f - 3);
    *force_mag = 6 * Emin * Rmini * f * (1 - f);
}

/* VdW_solver::calc_gensoft_factors (src/VdW_solver.cpp:754-771): Rmin is the cutoff. */
static void gensoft_factors(double mag_r, double Emin, double Rmin, double k0, double *force_mag, double *e) {
    const double r2 = mag_r * mag_r;
    const double R2 = Rmin * Rmin;
    const double g2 = r2 / R2, g3 = g2 * mag_r / Rmin;
    const double k0rm = k0 * Rmin, k0rm2 = k0rm * Rmin, eps = Emin / Rmin;
    *force_mag = 2 * (k0rm - 6 * eps) * g3 + 3 * (4 * eps - k0rm) * g2 + k0 * mag_r;
    const double emag = (0.5 * k0rm2 - 3 * Emin) * g2 * g2 + (4 * Emin - k0rm2) * g3 + 0.5 * k0 * r2 - Emin;
    *e = g_gw * g_gw * emag;
}

static void lj_factors(double mag_r, int k, int l, double Emin, double Rmin_6, double *force_mag, double *e);

/* The force magnitude and the energy of one Gauss-point pair at distance r, for the tests:
 * kind 1 lennard-jones, 2 ljsteric, 3 gensoft (its Rmin is the cutoff). */
void orc_ssint_factors(int kind, double r, double Emin, double Rmin, double k0, double *out) {
    const double R2 = Rmin * Rmin, R4 = R2 * R2, R6 = R4 * R2;
    if (kind == 3)
        gensoft_factors(r, Emin, Rmin, k0, &out[0], &out[1]);
    else if (kind == 2 && r < Rmin)
        lj_soft_factors(r, Emin, 1.0 / Rmin, &out[0], &out[1]);
    else
        lj_factors(r, 0, 0, Emin, R6, &out[0], &out[1]);
}

/* VdW_solver::calc_lj_factors (src/VdW_solver.cpp:713-731). */
static void lj_factors(double mag_r, int k, int l, double Emin, double Rmin_6, double *force_mag, double *e) {
    const double ri = 1. / mag_r;
    const double ri2 = ri * ri, ri4 = ri2 * ri2, ri6 = ri4 * ri2, ri7 = ri6 * ri;
    const double f6 = Rmin_6 * ri6;
    *force_mag = 12 * ri7 * Rmin_6 * Emin * (f6 - 1);
    *e = g_gw * g_gw * Emin * f6 * (f6 - 2);
    (void)k;
    (void)l;
}

/* VdW_solver::do_lj_interaction (src/VdW_solver.cpp:236-324) with calc_lj_force_pair_matrix
 * (:663-697): Gauss points of both faces, the force pair matrix (its lower triangle a copy of
 * the upper, as the reference fills it), the energy scaled by both areas, and the forces on
 * the three corners of each face added in corner order, face 1 then face 2. */
static void lj_pair(orc_blob *b1, int f1, orc_blob *b2, int f2, double *energy_out) {
    const int t1 = b1->ssint_type[f1], t2 = b2->ssint_type[f2];
    const double Emin = g_ss_emin[t1 * g_ss_ntypes + t2], Rmin = g_ss_rmin[t1 * g_ss_ntypes + t2];
    const double k0 = g_ss_k0[t1 * g_ss_ntypes + t2], Rmini = 1.0 / Rmin;
    const int *n1 = &b1->face_nodes[3 * f1], *n2 = &b2->face_nodes[3 * f2];
    double p[3][3], q[3][3], fpm[3][3][3];
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) {
            p[i][k] = g_geta[i][0] * b1->pos[3 * n1[0] + k] + g_geta[i][1] * b1->pos[3 * n1[1] + k] +
                      g_geta[i][2] * b1->pos[3 * n1[2] + k];
            q[i][k] = g_geta[i][0] * b2->pos[3 * n2[0] + k] + g_geta[i][1] * b2->pos[3 * n2[1] + k] +
                      g_geta[i][2] * b2->pos[3 * n2[2] + k];
        }
    const double Rmin_2 = Rmin * Rmin, Rmin_4 = Rmin_2 * Rmin_2, Rmin_6 = Rmin_4 * Rmin_2;
    double energy = 0.0, fm, e;
    for (int k = 0; k < 3; k++) {
        for (int l = k; l < 3; l++) {
            const double dx = p[k][0] - q[l][0], dy = p[k][1] - q[l][1], dz = p[k][2] - q[l][2];
            const double mag_r = sqrt(dx * dx + dy * dy + dz * dz);
            if (g_ss_kind == 3)
                gensoft_factors(mag_r, Emin, Rmin, k0, &fm, &e);
            else if (g_ss_kind == 2 && mag_r < Rmin)
                lj_soft_factors(mag_r, Emin, Rmini, &fm, &e);
            else
                lj_factors(mag_r, k, l, Emin, Rmin_6, &fm, &e);
            energy += l == k ? e : 2 * e;
            for (int c = 0; c < 3; c++) fpm[k][l][c] = fm * ((p[k][c] - q[l][c]) / mag_r);
            if (l != k)
                for (int c = 0; c < 3; c++) fpm[l][k][c] = fpm[k][l][c];
        }
    }
    const double ApAq = b1->face_geom[7 * f1 + 6] * b2->face_geom[7 * f2 + 6];
    *energy_out = energy * ApAq;
    for (int j = 0; j < 3; j++) {
        double F1[3] = {0, 0, 0}, F2[3] = {0, 0, 0};
        for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++) {
                const double c = g_gw * g_gw * g_geta[l][j];
                const double d = g_gw * g_gw * g_geta[k][j];
                for (int x = 0; x < 3; x++) F1[x] += c * fpm[k][l][x];
                for (int x = 0; x < 3; x++) F2[x] -= d * fpm[l][k][x];
            }
        for (int x = 0; x < 3; x++) b1->face_force[12 * f1 + 3 * j + x] += F1[x] * ApAq;
        for (int x = 0; x < 3; x++) b2->face_force[12 * f2 + 3 * j + x] += F2[x] * ApAq;
    }
}

/* ---- steric volume exclusion: overlap test, intersection volume, its gradient ------------ */

/* Steric parameters (SimulationParams steric_factor and steric_dr, src/SimulationParams.cpp:412-
 * 418, used as given, in meso units) and a flag for the geometry the reference cannot handle
 * (arr3Normalise2 throws on a zero vector, src/mat_vec_fns_II.cpp:164-177). */
static double g_st_factor = 1.0, g_st_dr = 5e-3;
static int g_st_throw;

void orc_world_set_steric(double factor, double dr) {
    g_st_factor = factor;
    g_st_dr = dr;
}

/* The vector helpers of src/mat_vec_fns_II.cpp, each rounding as the reference's loops do:
 * dot and norm sums start from 0.0 and add the terms in index order. */
static double vx_dot(const double *a, const double *b) { /* arr3arr3DotProduct :138-147 */
    double r = 0.0;
    r += a[0] * b[0];
    r += a[1] * b[1];
    r += a[2] * b[2];
    return r;
}
static void vx_sub(const double *a, const double *b, double *r) { /* arr3arr3Substract :117-122 */
    for (int i = 0; i < 3; i++) r[i] = a[i] - b[i];
}
static void vx_add(const double *a, const double *b, double *r) { /* arr3arr3Add :105-110 */
    for (int i = 0; i < 3; i++) r[i] = a[i] + b[i];
}
static void vx_neg(double *u) { /* arr3Resize(mOne, u) :179-184 */
    for (int i = 0; i < 3; i++) u[i] = -1.0 * u[i];
}
static void vx_cross(const double *u, const double *v, double *w) { /* arr3arr3VectorProduct :127-132 */
    w[0] = u[1] * v[2] - v[1] * u[2];
    w[1] = -u[0] * v[2] + v[0] * u[2];
    w[2] = u[0] * v[1] - v[0] * u[1];
}
static void vx_unit_cross(const double *u, const double *v, double *w) { /* getUnitNormal :311-322 */
    vx_cross(u, v, w);
    const double l = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (int i = 0; i < 3; i++) w[i] /= l;
}
static void vx_tangent(const double *a, const double *b, double *t) { /* tangent :295-306 */
    double w = 0;
    for (int i = 0; i < 3; i++) {
        t[i] = a[i] - b[i];
        w += t[i] * t[i];
    }
    w = sqrt(w);
    for (int i = 0; i < 3; i++) t[i] /= w;
}
static void vx_normalise_into(const double *e, double *n) { /* arr3Normalise2 :164-177 */
    double s = 0.0;
    for (int i = 0; i < 3; i++) s += e[i] * e[i];
    if (s == 0.0) g_st_throw = 1;
    s = sqrt(s);
    for (int i = 0; i < 3; i++) n[i] = e[i] / s;
}
static double vx_distance(const double *a, const double *b) { /* arr3arr3Distance :218-224 */
    double d = 0.0;
    for (int i = 0; i < 3; i++) d += (a[i] - b[i]) * (a[i] - b[i]);
    return sqrt(d);
}
static int vx_same_sign(double a, double b) { return a * b >= 0; } /* sameSign :34-38 */

/* getNormalInwards(f0, f1, f2, p3, n) (:361-377): the unit normal of the plane f0 f1 f2,
 * turned to the side of p3. */
static void vx_normal_inwards(const double *f0, const double *f1, const double *f2, const double *p3, double *n) {
    double pl1[3], pl2[3], aux[3];
    vx_sub(f1, f0, pl1);
    vx_sub(f2, f0, pl2);
    vx_cross(pl1, pl2, aux);
    vx_normalise_into(aux, n);
    vx_add(aux, f0, aux);
    const double d = -vx_dot(n, f0);
    const double t1 = vx_dot(p3, n) + d, t2 = vx_dot(aux, n) + d;
    if (!vx_same_sign(t1, t2)) vx_neg(n);
}

/* sameSidePlane (:383-398): vec and test on one side of the plane p1 p2 p3 (or on it). */
static int vx_same_side_plane(const double *vec, const double *test, const double *p1, const double *p2,
                              const double *p3) {
    double pl1[3], pl2[3], n[3];
    vx_sub(p2, p1, pl1);
    vx_sub(p3, p1, pl2);
    vx_cross(pl1, pl2, n);
    double s = 0.0;  /* arr3Normalise (:149-158), no zero test */
    for (int i = 0; i < 3; i++) s += n[i] * n[i];
    s = sqrt(s);
    for (int i = 0; i < 3; i++) n[i] /= s;
    const double d = -vx_dot(n, p1);
    return vx_same_sign(vx_dot(vec, n) + d, vx_dot(test, n) + d);
}

/* sameSideLine (:405-418): e and p1 on one side of the line p2 p3 within their plane. */
static int vx_same_side_line(const double *e, const double *p1, const double *p2, const double *p3) {
    double pl23[3], pl21[3], pl2e[3], v1[3], ve[3];
    vx_sub(p3, p2, pl23);
    vx_sub(p1, p2, pl21);
    vx_sub(e, p2, pl2e);
    vx_cross(pl21, pl23, v1);
    vx_cross(pl2e, pl23, ve);
    return vx_dot(v1, ve) >= 0;
}

/* nodeInTet (:434-441): on the inner side of all four face planes. */
static int vx_node_in_tet(const double *v, const double *t0, const double *t1, const double *t2, const double *t3) {
    return vx_same_side_plane(v, t0, t1, t2, t3) && vx_same_side_plane(v, t1, t2, t3, t0) &&
           vx_same_side_plane(v, t2, t3, t0, t1) && vx_same_side_plane(v, t3, t0, t1, t2);
}

/* intersectionPoint(ip, e1, e2, f1, f2, f3) (:573-588) with linePlaneIntersectionPoint
 * (:450-470) and isPointInFace (:531-538): where the edge e1 e2 crosses the triangle. */
static int vx_edge_face_point(double *ip, const double *e1, const double *e2, const double *f1, const double *f2,
                              const double *f3) {
    if (vx_same_side_plane(e1, e2, f1, f2, f3)) return 0;
    double v[3], pl1[3], pl2[3], pn[3];
    vx_sub(e2, e1, v);
    vx_sub(f2, f1, pl1);
    vx_sub(f3, f2, pl2);
    vx_unit_cross(pl1, pl2, pn);
    const double d = -vx_dot(pn, f1);
    const double t = -(vx_dot(pn, e1) + d) / vx_dot(pn, v);
    for (int i = 0; i < 3; i++) v[i] = t * v[i];
    vx_add(e1, v, ip);
    return vx_same_side_line(ip, f1, f2, f3) && vx_same_side_line(ip, f3, f1, f2) && vx_same_side_line(ip, f2, f3, f1);
}

/* getBAndN(f0, f1, f2, p3, t, b, n) (src/VolumeIntersection.cpp:179-204): for the edge leaving
 * f0 along t, b the inward unit normal of its face f0 f1 f2 and n the unit vector in that face
 * normal to t, towards f2. */
static void vx_face_frame(const double *f0, const double *f1, const double *f2, const double *p3, const double *t,
                          double *b, double *n) {
    double pl2[3], aux[3];
    vx_sub(f2, f0, pl2);
    vx_cross(t, pl2, aux);
    vx_normalise_into(aux, b);
    vx_add(aux, f0, aux);
    const double d = -vx_dot(b, f0);
    const double t1 = vx_dot(p3, b) + d, t2 = vx_dot(aux, b) + d;
    if (!vx_same_sign(t1, t2)) vx_neg(b);
    vx_unit_cross(b, t, n);
    vx_add(f0, n, aux);
    if (!vx_same_side_line(aux, f2, f1, f0)) vx_neg(n);
}

/* volumeForNode(tet0, tet1, tet2, tetN) (:259-298): vertex tetN's term of the polyhedron
 * volume sum, over its three edges and the two faces at each. */
static double vx_vertex_term(const double *a0, const double *a1, const double *a2, const double *aN) {
    const double *far[3] = {a0, a1, a2};
    static const int other[3][2] = {{1, 2}, {0, 2}, {0, 1}};
    double vol = 0.0, t[3], b[3], n[3];
    for (int e = 0; e < 3; e++) {
        vx_tangent(far[e], aN, t);
        const double pt = vx_dot(aN, t);
        for (int s = 0; s < 2; s++) {
            const double *q1 = far[other[e][s]], *q2 = far[other[e][1 - s]];
            vx_face_frame(aN, far[e], q1, q2, t, b, n);
            vol += pt * vx_dot(aN, n) * vx_dot(aN, b);
        }
    }
    return vol;
}

/* volumeForIntPointII (:547-655): the term of an intersection point ip of edge e1 e2 (of the
 * tetrahedron e1 e2 e3 e4) with the face f1 f2 f3 (of the tetrahedron f1 f2 f3 f4). */
static double vx_crossing_term(const double *ip, const double *e1, const double *e2, const double *e3,
                               const double *e4, const double *f1, const double *f2, const double *f3,
                               const double *f4) {
    double B[3][3], T[3][3], aux[3], n[3];
    vx_normal_inwards(f1, f2, f3, f4, B[0]);
    vx_tangent(e2, e1, T[0]);
    if (vx_same_side_plane(e1, f4, f1, f2, f3)) vx_neg(T[0]);
    vx_normal_inwards(e1, e2, e3, e4, B[1]);
    vx_unit_cross(B[0], B[1], T[1]);
    vx_add(ip, T[1], aux);
    if (!vx_same_side_plane(aux, e3, e4, e1, e2)) vx_neg(T[1]);
    vx_normal_inwards(e1, e2, e4, e3, B[2]);
    vx_unit_cross(B[0], B[2], T[2]);
    vx_add(ip, T[2], aux);
    if (!vx_same_side_plane(aux, e4, e3, e1, e2)) vx_neg(T[2]);
    double ipT[3], ipB[3];
    for (int i = 0; i < 3; i++) {
        ipT[i] = vx_dot(ip, T[i]);
        ipB[i] = vx_dot(ip, B[i]);
    }
    double vol = 0;
    /* the edge e1 e2 with the faces e1 e2 e3 and e1 e2 e4 */
    const double *side[2] = {e3, e4};
    for (int s = 0; s < 2; s++) {
        vx_unit_cross(B[1 + s], T[0], n);
        vx_add(ip, n, aux);
        if (!vx_same_side_line(aux, side[s], e2, e1)) vx_neg(n);
        vol += ipT[0] * vx_dot(ip, n) * ipB[1 + s];
    }
    /* the two new edges, each with the face f1 f2 f3 and its own face of e1 e2 e3 e4 */
    double V[3][3];
    vx_add(ip, T[1], V[0]);
    vx_add(ip, T[2], V[1]);
    for (int i = 0; i < 3; i++) V[2][i] = V[0][i];
    for (int i = 0; i < 2; i++) {
        vx_unit_cross(B[0], T[i + 1], n);
        vx_add(ip, n, aux);
        if (!vx_same_side_line(aux, V[i + 1], V[i], ip)) vx_neg(n);
        vol += ipT[i + 1] * vx_dot(ip, n) * ipB[0];
        vx_unit_cross(B[i + 1], T[i + 1], n);
        if (vx_dot(n, B[0]) < 0) vx_neg(n);
        vol += ipT[i + 1] * vx_dot(ip, n) * ipB[i + 1];
    }
    return vol;
}

/* The intersection polyhedron's vertices as they are found (at most 8 + 48), in the
 * reference's list W[56] and count ips. */
typedef struct {
    double w[56][3];
    int n;
} vx_points;

/* exists (:32-43): a point within 3 machine epsilons of one already listed. */
static int vx_listed(const vx_points *P, const double *p) {
    for (int i = 0; i < P->n; i++)
        if (vx_distance(p, P->w[i]) < 3.0 * DBL_EPSILON) return 1;
    return 0;
}

/* contribVolForIntersections (:929-949): edge ai aj of the tetrahedron ai aj a3 a4 against the
 * four faces of b0 b1 b2 b3, taken as (b0 b1 b2), (b1 b2 b3), (b2 b3 b0), (b3 b0 b1). */
static void vx_edge_terms(double *vol, vx_points *P, const double *ai, const double *aj, const double *a3,
                          const double *a4, const double *const *b) {
    double ip[3];
    for (int f = 0; f < 4; f++) {
        const double *q0 = b[f], *q1 = b[(f + 1) & 3], *q2 = b[(f + 2) & 3], *q3 = b[(f + 3) & 3];
        if (vx_edge_face_point(ip, ai, aj, q0, q1, q2) && !vx_listed(P, ip)) {
            memcpy(P->w[P->n++], ip, sizeof(ip));
            *vol += vx_crossing_term(ip, ai, aj, a3, a4, q0, q1, q2, q3);
        }
    }
}

/* volumeIntersectionII (:951-1024) with maxVolume (:48-66) and findCM (:95-104): the volume
 * common to tetrahedra A and B as -1/6 of the sum of the vertex terms of the intersection
 * polyhedron (vertices of either inside the other, then edge-face crossings, edge pairs in
 * the order (0,1), (0,2), (0,3), (1,2), (1,3), (2,3), A's edge before B's), zero for three
 * points or fewer, and replaced by 8 d^3 where it is negative or above that bound (d: the
 * distance from the points' mean to the last point listed, as the reference's loop leaves
 * it); cm, if asked for, the mean of the points. */
static double vx_intersection(const double *const *A, const double *const *B, int want_cm, double *cm) {
    vx_points P;
    P.n = 0;
    double vol = 0.0;
    static const int rest[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    for (int side = 0; side < 2; side++) {
        const double *const *X = side ? B : A, *const *Y = side ? A : B;
        for (int v = 0; v < 4; v++)
            if (!vx_listed(&P, X[v]) && vx_node_in_tet(X[v], Y[0], Y[1], Y[2], Y[3])) {
                memcpy(P.w[P.n++], X[v], 3 * sizeof(double));
                vol += vx_vertex_term(X[rest[v][0]], X[rest[v][1]], X[rest[v][2]], X[v]);
            }
    }
    static const int edge[6][4] = {{0, 1, 2, 3}, {0, 2, 1, 3}, {0, 3, 1, 2}, {1, 2, 0, 3}, {1, 3, 0, 2}, {2, 3, 0, 1}};
    for (int e = 0; e < 6; e++) {
        const int *k = edge[e];
        vx_edge_terms(&vol, &P, A[k[0]], A[k[1]], A[k[2]], A[k[3]], B);
        vx_edge_terms(&vol, &P, B[k[0]], B[k[1]], B[k[2]], B[k[3]], A);
    }
    vol *= -0.166666666666666667;
    if (P.n <= 3) return 0.;
    double c[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < P.n; i++) vx_add(c, P.w[i], c);
    const double s = 1. / P.n;
    for (int i = 0; i < 3; i++) c[i] = s * c[i];
    double d0 = 0.0, di = 0.0;
    for (int i = 0; i < P.n; i++) {
        di = vx_distance(c, P.w[i]);
        if (di > d0) d0 = di;
    }
    const double w = 8.0 * di * di * di;
    if (fabs(vol) > w || vol < 0) vol = w;
    if (want_cm) {
        for (int i = 0; i < 3; i++) cm[i] = 0.0;
        for (int i = 0; i < P.n; i++) vx_add(cm, P.w[i], cm);
        for (int i = 0; i < 3; i++) cm[i] = s * cm[i];
    }
    return vol;
}

/* getTetrahedraVolume (src/mat_vec_fns_II.cpp:708-714): six times the signed volume. */
static double vx_tet6(const double *p0, const double *p1, const double *p2, const double *p3) {
    return (p1[0] - p0[0]) * ((p1[1] - p2[1]) * (p2[2] - p3[2]) - (p2[1] - p3[1]) * (p1[2] - p2[2])) +
           (p2[0] - p1[0]) * ((p2[1] - p3[1]) * (p0[2] - p1[2]) - (p0[1] - p1[1]) * (p2[2] - p3[2])) +
           (p3[0] - p2[0]) * ((p0[1] - p1[1]) * (p1[2] - p2[2]) - (p1[1] - p2[1]) * (p0[2] - p1[2]));
}

/* getLocalCoordinatesForLinTet (:725-733): the barycentric coordinates of p in t0..t3. */
static void vx_local_coords(const double *const *t, const double *p, double *phi) {
    phi[0] = vx_tet6(p, t[1], t[2], t[3]);
    phi[1] = -vx_tet6(p, t[0], t[2], t[3]);
    phi[2] = vx_tet6(p, t[0], t[1], t[3]);
    phi[3] = -vx_tet6(p, t[0], t[1], t[2]);
    const double v = vx_tet6(t[0], t[1], t[2], t[3]);
    for (int i = 0; i < 4; i++) phi[i] /= v;
}

/* tet_a_tetII (src/CheckTetrahedraOverlap.cpp:208-337): the separating-axis overlap test of
 * Ganovelli, Ponchio and Rocchini (JGT 2002) -- A's four face planes (with the edge planes at
 * each pair of them), then, unless some vertex of B lies inside all of A's faces, B's four. */
static int vx_side_mask(const double *c) {
    int m = 0;
    for (int k = 0; k < 4; k++)
        if (c[k] > 0.0) m |= 1 << k;
    return m;
}
static int vx_edge_separates(const double *c0, const double *c1, int m0, int m1) {
    if ((m0 | m1) != 15) return 0;
    m0 &= (m0 ^ m1);
    m1 &= (m0 ^ m1);
    static const int pr[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (int q = 0; q < 6; q++) {
        const int i = pr[q][0], j = pr[q][1];
        const double det = c0[j] * c1[i] - c0[i] * c1[j];
        if ((m0 & (1 << i)) && (m1 & (1 << j)) && det > 0.0) return 0;
        if ((m0 & (1 << j)) && (m1 & (1 << i)) && det < 0.0) return 0;
    }
    return 1;
}
static int vx_tets_overlap(const double *const *A, const double *const *B) {
    double PA[4][3], e[5][3], n[3], C[4][4];
    int mask[4];
    for (int k = 0; k < 4; k++) vx_sub(B[k], A[0], PA[k]);
    vx_sub(A[1], A[0], e[0]);
    vx_sub(A[2], A[0], e[1]);
    vx_sub(A[3], A[0], e[2]);
    /* A's faces 0..2 through A0, then face 3 through A1 */
    for (int f = 0; f < 4; f++) {
        if (f == 0) vx_cross(e[0], e[1], n);
        else if (f == 1) vx_cross(e[2], e[0], n);
        else if (f == 2) vx_cross(e[1], e[2], n);
        else {
            vx_sub(A[3], A[1], e[4]);
            vx_sub(A[2], A[1], e[3]);
            vx_cross(e[4], e[3], n);
        }
        for (int k = 0; k < 4; k++) {
            if (f < 3)
                C[f][k] = PA[k][0] * n[0] + PA[k][1] * n[1] + PA[k][2] * n[2];
            else
                C[f][k] = (B[k][0] - A[1][0]) * n[0] + (B[k][1] - A[1][1]) * n[1] + (B[k][2] - A[1][2]) * n[2];
        }
        mask[f] = vx_side_mask(C[f]);
        if (mask[f] == 15) return 0;
        for (int g = 0; g < f; g++)
            if (vx_edge_separates(C[g], C[f], mask[g], mask[f])) return 0;
    }
    if ((mask[0] | mask[1] | mask[2] | mask[3]) != 15) return 1;
    double PB[4][3], d[5][3];
    for (int k = 0; k < 4; k++) vx_sub(A[k], B[0], PB[k]);
    vx_sub(B[1], B[0], d[0]);
    vx_sub(B[2], B[0], d[1]);
    for (int f = 0; f < 4; f++) {
        if (f == 0) vx_cross(d[0], d[1], n);
        else if (f == 1) {
            vx_sub(B[3], B[0], d[2]);
            vx_cross(d[2], d[0], n);
        } else if (f == 2) vx_cross(d[1], d[2], n);
        else {
            vx_sub(B[3], B[1], d[4]);
            vx_sub(B[2], B[1], d[3]);
            vx_cross(d[4], d[3], n);
        }
        int all = 1;
        for (int k = 0; k < 4 && all; k++) {
            const double s = f < 3 ? PB[k][0] * n[0] + PB[k][1] * n[1] + PB[k][2] * n[2]
                                   : (A[k][0] - B[1][0]) * n[0] + (A[k][1] - B[1][1]) * n[1] + (A[k][2] - B[1][2]) * n[2];
            all = s > 0.0;
        }
        if (all) return 0;
    }
    return 1;
}

/* Face::getTetraIntersectionVolumeTotalGradientAndShapeFunctions (src/Face.cpp:473-514) and
 * VdW_solver::do_steric_interaction (src/VdW_solver.cpp:458-487) for the tetrahedra A (face 1:
 * its corners and opposite vertex) and B (face 2): the common volume V with the points' mean
 * cm, the barycentric coordinates of cm in each, and the one-sided differences
 *   dV/dr_d = (V(A, B + dr e_d) - V) / dr - (V(B, A + dr e_d) - V) / dr;
 * then V and the gradient scaled by the steric factor, V the pair's energy, phi1_j dV/dr the
 * force on A's node j and -phi2_j dV/dr on B's.  Returns 0 (nothing applied) where V is 0. */
static int steric_forces(const double *const *A, const double *const *B, double dr, double factor, double *energy,
                         double *fA, double *fB) {
    double cm[3], phi1[4], phi2[4], g[3];
    double vol = vx_intersection(A, B, 1, cm);
    if (vol == 0) return 0;
    vx_local_coords(A, cm, phi1);
    vx_local_coords(B, cm, phi2);
    for (int dir = 0; dir < 3; dir++) {
        double dx[3] = {0, 0, 0}, sA[4][3], sB[4][3];
        dx[dir] = dr;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 3; j++) {
                sA[i][j] = A[i][j] + dx[j];
                sB[i][j] = B[i][j] + dx[j];
            }
        const double *SA[4] = {sA[0], sA[1], sA[2], sA[3]}, *SB[4] = {sB[0], sB[1], sB[2], sB[3]};
        double vm = vx_intersection(A, SB, 0, NULL);
        g[dir] = (vm - vol) / dr;
        vm = vx_intersection(B, SA, 0, NULL);
        g[dir] -= (vm - vol) / dr;
    }
    vol *= factor;
    for (int i = 0; i < 3; i++) g[i] = factor * g[i];
    *energy = vol;
    for (int j = 0; j < 4; j++)
        for (int x = 0; x < 3; x++) {
            fA[3 * j + x] = phi1[j] * g[x];
            fB[3 * j + x] = (-1.0 * phi2[j]) * g[x];
        }
    return 1;
}

/* Face 1 of blob b1 against face 2 of blob b2 (steric, ljsteric and gensoft): the overlap test
 * of Face::checkTetraIntersection (src/Face.cpp:399-404), then the steric forces added to both
 * faces' Face::force[0..3]; returns 1 where they applied. */
static int steric_pair(orc_blob *b1, int f1, orc_blob *b2, int f2, double *energy) {
    const double *A[4], *B[4];
    for (int k = 0; k < 3; k++) {
        A[k] = &b1->pos[3 * b1->face_nodes[3 * f1 + k]];
        B[k] = &b2->pos[3 * b2->face_nodes[3 * f2 + k]];
    }
    if (b1->ssint_opp[f1] < 0 || b2->ssint_opp[f2] < 0) { /* Face::n[3] is always set in the reference */
        g_st_throw = 2;
        return 0;
    }
    A[3] = &b1->pos[3 * b1->ssint_opp[f1]];
    B[3] = &b2->pos[3 * b2->ssint_opp[f2]];
    if (!vx_tets_overlap(A, B)) return 0;
    double fA[12], fB[12];
    if (!steric_forces(A, B, g_st_dr, g_st_factor, energy, fA, fB)) return 0;
    for (int k = 0; k < 12; k++) {
        b1->face_force[12 * f1 + k] += fA[k];
        b2->face_force[12 * f2 + k] += fB[k];
    }
    return 1;
}

/* Test hooks: tetrahedra as [4][3] vertex arrays. */
static void tet_ptrs(const double *t, const double **p) {
    for (int k = 0; k < 4; k++) p[k] = t + 3 * k;
}
int orc_tets_overlap(const double *a, const double *b) {
    const double *A[4], *B[4];
    tet_ptrs(a, A);
    tet_ptrs(b, B);
    return vx_tets_overlap(A, B);
}
double orc_tet_intersection(const double *a, const double *b, int want_cm, double *cm) {
    const double *A[4], *B[4];
    tet_ptrs(a, A);
    tet_ptrs(b, B);
    return vx_intersection(A, B, want_cm, cm);
}
void orc_tet_local_coords(const double *t, const double *p, double *phi) {
    const double *T[4];
    tet_ptrs(t, T);
    vx_local_coords(T, p, phi);
}
/* out: [0] the energy, [1..12] A's node forces, [13..24] B's; returns whether they apply. */
int orc_steric_forces(const double *a, const double *b, double dr, double factor, double *out) {
    const double *A[4], *B[4];
    tet_ptrs(a, A);
    tet_ptrs(b, B);
    g_st_throw = 0;
    const int r = steric_forces(A, B, dr, factor, &out[0], &out[1], &out[13]);
    return g_st_throw ? -1 : r;
}

/* consider_interaction's further tests for the types other than lennard-jones
 * (src/VdW_solver.cpp:354-381): the faces must be in front of each other (sep from face i's
 * centroid to face j's, arr3arr3DotProduct: sep.n_i < 0 with sep.n_j > 0 rejects), and two
 * faces of one blob may share no node (the three corners and the opposite vertex). */
static int ssint_steric_checks(const orc_blob *b1, int f1, const orc_blob *b2, int f2, int same_blob) {
    const double *c1 = &b1->face_geom[7 * f1], *c2 = &b2->face_geom[7 * f2];
    double sep[3], d1 = 0.0, d2 = 0.0;
    for (int k = 0; k < 3; k++) sep[k] = c2[k] - c1[k];
    for (int k = 0; k < 3; k++) d1 += sep[k] * c1[3 + k];
    for (int k = 0; k < 3; k++) d2 += sep[k] * c2[3 + k];
    if (d1 < 0 && d2 > 0) return 0;
    if (g_ss_inc_self && same_blob) {
        if (b1->ssint_opp[f1] == b2->ssint_opp[f2]) return 0;
        int n1[4] = {b1->face_nodes[3 * f1], b1->face_nodes[3 * f1 + 1], b1->face_nodes[3 * f1 + 2], b1->ssint_opp[f1]};
        int n2[4] = {b2->face_nodes[3 * f2], b2->face_nodes[3 * f2 + 1], b2->face_nodes[3 * f2 + 2], b2->ssint_opp[f2]};
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
                if (n2[j] == n1[i]) return 0;
    }
    return 1;
}

/* VdW_solver::solve (src/VdW_solver.cpp:158-210) with consider_interaction (:332-384): each
 * pool entry i against the stacks of its cell and the 26 around it (top down), pairs with
 * i < j, of different blobs unless inc_self_ssint, whose normals do not face the same way (dot
 * product not above zero), and for the types other than lennard-jones the further tests. */
static void ssint_solve(orc_blob **blobs, int nblobs) {
    free(g_ss_energy);
    g_ss_energy = xcalloc((size_t)nblobs * nblobs, sizeof(double));
    g_ss_nblobs = nblobs;
    for (int i = 0; i < g_ss_npool; i++) {
        const int bi = g_ss_pool_blob[i], fi = g_ss_pool_face[i], ci = g_ss_pool_cell[i];
        const int cx = ci / (g_ss_N[1] * g_ss_N[2]), cy = (ci / g_ss_N[2]) % g_ss_N[1], cz = ci % g_ss_N[2];
        const double *ni = &blobs[bi]->face_geom[7 * fi + 3];
        for (int c = 0; c < 27; c++) {
            int x[3] = {cx + g_adj[c][0], cy + g_adj[c][1], cz + g_adj[c][2]};
            cell_pbc(x);
            for (int j = g_ss_root[x[0] * g_ss_N[1] * g_ss_N[2] + x[1] * g_ss_N[2] + x[2]]; j >= 0; j = g_ss_next[j]) {
                if (!(i < j)) continue;
                const int bj = g_ss_pool_blob[j], fj = g_ss_pool_face[j];
                if (!g_ss_inc_self && bi == bj) continue;
                const double *nj = &blobs[bj]->face_geom[7 * fj + 3];
                if (ni[0] * nj[0] + ni[1] * nj[1] + ni[2] * nj[2] > 0.0) continue;
                if (g_ss_kind != 1 && !ssint_steric_checks(blobs[bi], fi, blobs[bj], fj, bi == bj)) continue;
                /* Steric_solver, LJSteric_solver and GenSoftSSINT_solver::do_interaction
                 * (src/Steric_solver.cpp:28-38, src/LJSteric_solver.cpp:34-52,
                 * src/GenSoftSSINT_solver.cpp:29-47): the steric forces where the face elements
                 * overlap with a nonzero volume, otherwise the pair potential (none for steric) */
                double e;
                if (g_ss_kind != 1 && steric_pair(blobs[bi], fi, blobs[bj], fj, &e))
                    g_ss_energy[bi * nblobs + bj] += e;
                else if (g_ss_kind != 4) {
                    lj_pair(blobs[bi], fi, blobs[bj], fj, &e);
                    g_ss_energy[bi * nblobs + bj] += e;
                }
            }
        }
    }
}

/* VdW_solver::get_field_energy(-1, -1) (src/VdW_solver.cpp:635-661): every blob pair's sum. */
double orc_world_ssint_energy(void) {
    double e = 0.0;
    for (int i = 0; i < g_ss_nblobs; i++)
        for (int j = 0; j < g_ss_nblobs; j++) e += g_ss_energy[i * g_ss_nblobs + j];
    return e;
}

int orc_world_step(orc_blob **blobs, int nblobs, orc_stream *g) {
    int inv = 0, err = 0;
    for (int i = 0; i < nblobs; i++) {
        blob_zero_forces(blobs[i]);
        blob_wrap(blobs[i]); /* src/World.cpp:1843-1904 */
    }
    if (g_ss_ntypes > 0) { /* the es_update refresh, at the top of the step (src/World.cpp:1821-1925) */
        if (g_ss_es_count == g_ss_es_update) {
            g_ss_es_count = 1;
            if (ssint_refresh(blobs, nblobs) != 0) return -2;
        } else {
            g_ss_es_count++;
        }
    }
    world_apply_springs(blobs); /* src/World.cpp:1933 */
    if (g_ss_ntypes > 0) { /* src/World.cpp:2002-2004 */
        g_st_throw = 0;
        ssint_solve(blobs, nblobs);
        if (g_st_throw) return g_st_throw == 2 ? -4 : -3; /* -3: a zero vector normalised, where the
                                                          * reference throws; -4: a face without its
                                                          * opposite vertex */
    }
    for (int i = 0; i < nblobs; i++) inv += orc_blob_update_internal_forces(blobs[i], g);
    if (inv > 0) return inv;
    for (int i = 0; i < nblobs; i++)
        if (orc_blob_update_positions(blobs[i], g) != 0) err = 1;
    return err ? -1 : 0;
}

/* ---- element/assembly/solver probes for tests ---------------------------- */

/* Element phase + aggregation only (no solve): fills node_force, dpsi, vol, visc and
 * leaves the linear-system RHS in out_rhs [3*N_t]. */
int orc_blob_forces_only(orc_blob *b, orc_stream *g, double *out_rhs) {
    blob_zero_forces(b);
    int inv = orc_blob_update_internal_forces(b, g);
    blob_aggregate_forces(b, g);
    memcpy(out_rhs, b->force, sizeof(double) * 3 * (size_t)b->num_nodes);
    return inv;
}

/* orc_blob_forces_only with the element phase drawing from g_elem and the force gather's
 * Stokes noise from g_node: blob k of a World whose blobs share one stream draws its element
 * noise after the earlier blobs' element draws and its node noise after every blob's element
 * draws and the earlier blobs' node draws (World::run, src/World.cpp:2020-2041). */
int orc_blob_forces_split(orc_blob *b, orc_stream *g_elem, orc_stream *g_node, double *out_rhs) {
    blob_zero_forces(b);
    int inv = orc_blob_update_internal_forces(b, g_elem);
    blob_aggregate_forces(b, g_node);
    memcpy(out_rhs, b->force, sizeof(double) * 3 * (size_t)b->num_nodes);
    return inv;
}

/* The blob's solver on a caller RHS (in/out); NoMassCGSolver with the element matrices
 * currently held. */
int orc_blob_solve(orc_blob *b, double *x) { return blob_solver_solve(b, x); }

/* V * in with the matrix as last built (call after a solve or orc_blob_build). */
void orc_blob_build(orc_blob *b) { blob_build_matrix(b); }
void orc_blob_apply_matrix(const orc_blob *b, const double *in, double *out) { blob_spmv(b, in, out); }

int orc_blob_nnz(const orc_blob *b) { return b->nnz; }
int orc_blob_num_rows(const orc_blob *b) { return b->num_rows; }
void orc_blob_get_csr(const orc_blob *b, int *key, int *col, double *val) {
    memcpy(key, b->key, sizeof(int) * ((size_t)b->num_rows + 1));
    memcpy(col, b->col, sizeof(int) * (size_t)b->nnz);
    memcpy(val, b->val, sizeof(double) * (size_t)b->nnz);
}
int orc_blob_last_cg_iterations(const orc_blob *b) { return b->last_cg_iterations; }

/* Copy out arrays by name for golden comparisons. */
void orc_blob_get(const orc_blob *b, const char *what, double *out) {
    size_t N = b->num_nodes, E = b->num_elements;
    if (!strcmp(what, "pos")) memcpy(out, b->pos, sizeof(double) * 3 * N);
    else if (!strcmp(what, "vel")) memcpy(out, b->vel, sizeof(double) * 3 * N);
    else if (!strcmp(what, "force")) memcpy(out, b->force, sizeof(double) * 3 * N);
    else if (!strcmp(what, "rhs")) memcpy(out, b->f, sizeof(double) * 3 * N);
    else if (!strcmp(what, "dpsi")) memcpy(out, b->dpsi, sizeof(double) * 12 * E);
    else if (!strcmp(what, "vol")) memcpy(out, b->vol, sizeof(double) * E);
    else if (!strcmp(what, "vol_0")) memcpy(out, b->vol_0, sizeof(double) * E);
    else if (!strcmp(what, "J_inv_0")) memcpy(out, b->J_inv_0, sizeof(double) * 9 * E);
    else if (!strcmp(what, "F_ij")) memcpy(out, b->F_ij, sizeof(double) * 9 * E);
    else if (!strcmp(what, "visc")) memcpy(out, b->visc, sizeof(double) * 144 * E);
    else if (!strcmp(what, "node_force")) memcpy(out, b->node_force, sizeof(double) * 30 * E);
    else if (!strcmp(what, "stress_mag")) memcpy(out, b->stress_mag, sizeof(double) * E);
    else if (!strcmp(what, "stokes_drag")) memcpy(out, b->stokes_drag, sizeof(double) * N);
    else if (!strcmp(what, "last_det")) memcpy(out, b->last_det, sizeof(double) * E);
    else if (!strcmp(what, "face_force") && b->face_force)
        memcpy(out, b->face_force, sizeof(double) * 12 * (size_t)b->n_faces);
    else if (!strcmp(what, "material")) {
        for (size_t e = 0; e < E; e++) {
            out[5 * e] = b->rho[e];
            out[5 * e + 1] = b->A[e];
            out[5 * e + 2] = b->B[e];
            out[5 * e + 3] = b->G[e];
            out[5 * e + 4] = b->E[e];
        }
    }
}

/* Strain energy of Blob::make_measurements (src/Blob.cpp:1424-1432), from F_ij of the
 * last element update: 1/2 sum vol0 [G(F:F - 3) + 1/2 C (J^2 - 1) - (C + 2G) ln J]. */
double orc_blob_strain_energy(const orc_blob *b) {
    double se = 0.0;
    for (int n = 0; n < b->num_elements; n++) {
        const double *F = &b->F_ij[9 * n];
        /* mat3_double_contraction (src/mat_vec_fns.cpp:139-143) order */
        double dc = F[0] * F[0] + F[4] * F[4] + F[8] * F[8] + F[1] * F[1] + F[3] * F[3] + F[2] * F[2] +
                    F[6] * F[6] + F[5] * F[5] + F[7] * F[7];
        double C = b->E[n] - (2.0 / 3.0) * b->G[n];
        double t = b->vol[n] / b->vol_0[n];
        se += b->vol_0[n] * (b->G[n] * (dc - 3) + 0.5 * C * (t * t - 1) - (C + 2 * b->G[n]) * log(t));
    }
    return 0.5 * se;
}

/* For the tests' intersection counts (the squidgy-spheres check of the reference's
 * tests/physics/squidgy_steric/checkSquidgySpheres.cpp): the number of pairs of tetrahedra, one
 * from each list ([n][4][3] vertex arrays), with a nonzero common volume.  Pairs whose bounding
 * boxes are apart have none and are skipped. */
long orc_count_tet_intersections(int na, const double *ta, int nb, const double *tb) {
    long n = 0;
    for (int i = 0; i < na; i++) {
        const double *a = ta + 12 * (size_t)i;
        double alo[3], ahi[3];
        for (int k = 0; k < 3; k++) {
            alo[k] = fmin(fmin(a[k], a[3 + k]), fmin(a[6 + k], a[9 + k]));
            ahi[k] = fmax(fmax(a[k], a[3 + k]), fmax(a[6 + k], a[9 + k]));
        }
        for (int j = 0; j < nb; j++) {
            const double *b = tb + 12 * (size_t)j;
            int apart = 0;
            for (int k = 0; k < 3 && !apart; k++) {
                const double lo = fmin(fmin(b[k], b[3 + k]), fmin(b[6 + k], b[9 + k]));
                const double hi = fmax(fmax(b[k], b[3 + k]), fmax(b[6 + k], b[9 + k]));
                apart = hi < alo[k] || lo > ahi[k];
            }
            if (!apart && orc_tet_intersection(a, b, 0, NULL) != 0) n++;
        }
    }
    return n;
}
