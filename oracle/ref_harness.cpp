/*
 * ref_harness.cpp -- drives the reference's OWN compiled hot-path objects
 * (tetra_element_linear, NoMassCGSolver, ConjugateGradientSolver, MassLumpedSolver,
 * SparsityPattern, SparseMatrixFixedPattern, SparseMatrixTypes, mat_vec_fns, mat_vec_fns_II,
 * RngStream, mesh_node, VolumeIntersection, CheckTetrahedraOverlap) from /root/reference/src.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the checker).  Built by oracle/Makefile
 * into oracle/_ref/libffea_ref.so; nothing in the product links it.
 *
 * Blob.cpp / World.cpp cannot be compiled here (they include Eigen, which the
 * reference's build downloads), so the short Blob-level loops that call the
 * element and solver objects are restated below, each citing the lines it follows.
 * Every numerical kernel (element math, pattern, matrix build, SpMV, CG, RNG) is the
 * reference's own code.  The C API mirrors oracle/ffea_oracle.c (ref_* for orc_*), so
 * the tests run one scenario through both and compare bit for bit.
 */
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <set>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "CheckTetrahedraOverlap.h"
#include "RngStream.h"
#include "VolumeIntersection.h"
#include "mat_vec_fns.h"
#include "mat_vec_fns_II.h"
#include "mesh_node.h"
#include "tetra_element_linear.h"
#define private public /* read NoMassCGSolver::V for matrix comparisons */
#include "NoMassCGSolver.h"
#undef private
#include "ConjugateGradientSolver.h"
#include "MassLumpedSolver.h"

namespace {
const double meso_volume = 4.913e-30, meso_mass = 1.994307387553024e-26;
const double meso_pressure = 8.4305832688784826e8, meso_time = 3.7302915560886126e-13;
}  // namespace

struct ref_blob {
    int N, E, num_pinned;
    mesh_node *node;
    tetra_element_linear *elem;
    NoMassCGSolver *solver;
    SimulationParams *params;
    unsigned char *params_storage;
    vector3 *force;
    int *pinned;
    double *rhs;
    int last_inversions;
    int linear_solver;  // FFEA_ITERATIVE_SOLVER, FFEA_MASSLUMPED_SOLVER or FFEA_NOMASS_CG_SOLVER
    Solver *msolver;    // the mass-ful solver when one is selected
};

extern "C" {

/* ---- RngStream ---------------------------------------------------------- */
void *ref_rng_new(void) { return new RngStream[1]; }
/* one stream per OpenMP thread, as World::init allocates them (src/World.cpp:313) */
void *ref_rng_new_array(int n) { return new RngStream[n]; }
int ref_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
/* One rotational constant force on a node at p, as Blob::apply_ctforces CTF 2 computes it
 * (src/Blob.cpp:2104-2136, type 'p'): the reference's own intersectingPointToLine,
 * arr3arr3Substract, arr3arr3VectorProduct, mag and arr3Resize (src/mat_vec_fns_II.cpp). */
void ref_rot_force(const double *p, const double *axis, double F, int torque, double *out) {
    vector3 pos;
    pos.assign(p[0], p[1], p[2]);
    arr3 r, f;
    double a[6];
    for (int k = 0; k < 6; k++) a[k] = axis[k];
    intersectingPointToLine<scalar, arr3>(pos, arr3_view<scalar, arr3>(a, 3), arr3_view<scalar, arr3>(a + 3, 3), r);
    arr3arr3Substract<scalar, arr3>(r, pos.data, r);
    arr3arr3VectorProduct<scalar, arr3>(r, arr3_view<scalar, arr3>(a + 3, 3), f);
    if (torque) {
        scalar rsize = mag<scalar, arr3>(r);
        arr3Resize<scalar, arr3>(F / rsize, f);
    } else {
        arr3Resize<scalar, arr3>(F, f);
    }
    for (int k = 0; k < 3; k++) out[k] = f[k];
}

void ref_rng_delete(void *g) { delete[] static_cast<RngStream *>(g); }
int ref_rng_set_package_seed(const unsigned long seed[6]) { return RngStream::SetPackageSeed(seed) ? 1 : 0; }
double ref_rng_randu01(void *g) { return static_cast<RngStream *>(g)->RandU01(); }
void ref_rng_get_state(void *g, unsigned long out[6]) { static_cast<RngStream *>(g)->GetState(out); }
int ref_rng_set_seed(void *g, const unsigned long seed[6]) { return static_cast<RngStream *>(g)->SetSeed(seed) ? 1 : 0; }
void ref_rng_advance_state(void *g, long e, long c) { static_cast<RngStream *>(g)->AdvanceState(e, c); }

/* World::init RNG.2 (src/World.cpp:299-305) with the host libc rand().  The reference's int
 * sum rand()+rand() is undefined when it exceeds INT_MAX; its optimised build adds in 64 bits
 * (the type of twoMax), which is what its checkpoint fixture cyl_160_fine.fcp holds. */
void ref_package_seed_from_rng_seed(int rng_seed, unsigned long sixseed[6]) {
    unsigned long twoMax[2] = {4294967087UL, 4294944443UL};
    srand(rng_seed);
    for (int i = 0; i < 6; i++) {
        unsigned long r1 = (unsigned long)rand();
        unsigned long r2 = (unsigned long)rand();
        sixseed[i] = (r1 + r2) % twoMax[i / 3];
    }
}

/* ---- Blob --------------------------------------------------------------- */
struct ref_blob;
/* The blob's reference objects, for drivers that hand them to a Solver (mesh_node[N],
 * tetra_element_linear[E], SimulationParams). */
void *ref_blob_node_array(ref_blob *b);
void *ref_blob_element_array(ref_blob *b);
void *ref_blob_params_ptr(ref_blob *b);
void ref_blob_set_max_iterations_cg(ref_blob *b, int n);
ref_blob *ref_blob_create(int num_nodes, int num_elements, const double *pos, const int *topo,
                          const double *mat_si, const double *stokes_radius, int num_pinned,
                          const int *pinned, double dt, double kT, double epsilon2, int max_iterations_cg,
                          int calc_noise, int calc_stokes, double stokes_visc) {
    ref_blob *b = new ref_blob();
    b->N = num_nodes;
    b->E = num_elements;
    /* SimulationParams.o cannot be linked (its TU needs a generated header); the solver
     * and element objects only read these plain fields, so a zeroed instance suffices. */
    b->params_storage = new unsigned char[sizeof(SimulationParams)]();
    b->params = reinterpret_cast<SimulationParams *>(b->params_storage);
    b->params->dt = dt;
    b->params->kT = kT;
    b->params->epsilon2 = epsilon2;
    b->params->max_iterations_cg = max_iterations_cg;
    b->params->calc_noise = calc_noise;
    b->params->calc_stokes = calc_stokes;
    b->params->stokes_visc = stokes_visc;
    b->params->calc_es = 0;

    /* Blob::load_nodes (src/Blob.cpp:2480-2495) */
    b->node = new mesh_node[num_nodes];
    for (int i = 0; i < num_nodes; i++) {
        b->node[i].pos.x = pos[3 * i];
        b->node[i].pos.y = pos[3 * i + 1];
        b->node[i].pos.z = pos[3 * i + 2];
        b->node[i].vel.x = b->node[i].vel.y = b->node[i].vel.z = 0;
        b->node[i].index = i;
        b->node[i].stokes_radius = stokes_radius[i];
    }
    /* Blob::load_topology (src/Blob.cpp:2606-2622) + load_material_params (:2861-2866) */
    b->elem = new tetra_element_linear[num_elements];
    for (int e = 0; e < num_elements; e++) {
        for (int j = 0; j < 10; j++) b->elem[e].n[j] = &b->node[topo[10 * e + j]];
        for (int j = 0; j < 4; j++) b->node[topo[10 * e + j]].set_linear();
        b->elem[e].index = e;
        const double *m = &mat_si[6 * e];
        b->elem[e].rho = m[0] * meso_volume / meso_mass;
        b->elem[e].A = m[1] / (meso_pressure * meso_time);
        b->elem[e].B = m[2] / (meso_pressure * meso_time) - (2.0 / 3.0) * m[1];
        b->elem[e].G = m[3] / meso_pressure;
        b->elem[e].E = m[4] / meso_pressure;
        b->elem[e].dielectric = m[5];
    }
    b->num_pinned = num_pinned;
    b->pinned = new int[num_pinned > 0 ? num_pinned : 1];
    for (int i = 0; i < num_pinned; i++) b->pinned[i] = pinned[i];

    /* Blob::init order (src/Blob.cpp:335-412) */
    for (int e = 0; e < num_elements; e++) b->elem[e].linearise_element();
    /* calc_rest_state_info (src/Blob.cpp:3768-3794) */
    for (int e = 0; e < num_elements; e++) {
        matrix3 J;
        scalar temp;
        b->elem[e].calculate_jacobian(J);
        mat3_invert(J, b->elem[e].J_inv_0, &temp);
        b->elem[e].calc_shape_function_derivatives_and_volume(J);
        b->elem[e].vol_0 = b->elem[e].vol;
        b->elem[e].mass = b->elem[e].vol_0 * b->elem[e].rho;
    }
    /* calculate_node_element_connectivity (src/Blob.cpp:3974-4019) */
    for (int i = 0; i < num_nodes; i++) b->node[i].num_element_contributors = 0;
    for (int e = 0; e < num_elements; e++)
        for (int j = 0; j < 10; j++) b->node[b->elem[e].n[j]->index].num_element_contributors++;
    for (int i = 0; i < num_nodes; i++)
        b->node[i].force_contributions = new vector3 *[b->node[i].num_element_contributors + 1];
    int *cnt = new int[num_nodes]();
    for (int e = 0; e < num_elements; e++)
        for (int j = 0; j < 10; j++) {
            int ni = b->elem[e].n[j]->index;
            b->node[ni].force_contributions[cnt[ni]++] = &b->elem[e].node_force[j];
        }
    delete[] cnt;
    b->solver = new NoMassCGSolver();
    std::set<int> bsite;
    b->solver->init(num_nodes, num_elements, b->node, b->elem, b->params, num_pinned, b->pinned, bsite);
    b->force = new vector3[num_nodes];
    for (int i = 0; i < num_nodes; i++) b->force[i].x = b->force[i].y = b->force[i].z = 0;
    if (calc_stokes == 1)
        for (int i = 0; i < num_nodes; ++i)
            b->node[i].stokes_drag = 6.0 * ffea_const::pi * stokes_visc * b->node[i].stokes_radius;
    b->rhs = new double[3 * (size_t)num_nodes]();
    b->linear_solver = FFEA_NOMASS_CG_SOLVER;
    b->msolver = nullptr;
    return b;
}

/* Blob::init's solver choice (src/Blob.cpp:358-380) for the two mass-ful solvers. */
int ref_blob_set_solver(ref_blob *b, int solver) {
    if (solver != FFEA_ITERATIVE_SOLVER && solver != FFEA_MASSLUMPED_SOLVER && solver != FFEA_NOMASS_CG_SOLVER)
        return -1;
    delete b->msolver;
    b->msolver = nullptr;
    b->linear_solver = solver;
    if (solver == FFEA_NOMASS_CG_SOLVER) return 0;
    if (solver == FFEA_ITERATIVE_SOLVER)
        b->msolver = new ConjugateGradientSolver();
    else
        b->msolver = new MassLumpedSolver();
    std::set<int> bsite;
    return b->msolver->init(b->N, b->E, b->node, b->elem, b->params, b->num_pinned, b->pinned, bsite) == FFEA_OK ? 0 : -1;
}

void ref_blob_set_velocities(ref_blob *b, const double *vel) {
    for (int i = 0; i < b->N; i++) {
        b->node[i].vel.x = vel[3 * i];
        b->node[i].vel.y = vel[3 * i + 1];
        b->node[i].vel.z = vel[3 * i + 2];
    }
}

/* Kinetic energy of Blob::make_measurements (src/Blob.cpp:1385-1407, :1431). */
double ref_blob_kinetic_energy(ref_blob *b) {
    if (b->linear_solver == FFEA_NOMASS_CG_SOLVER) return 0.0;
    scalar kenergy = 0.0;
    vector12 vec;
    for (int n = 0; n < b->E; n++) {
        tetra_element_linear &e = b->elem[n];
        e.get_element_velocity_vector(vec);
        e.apply_element_mass_matrix(vec);
        kenergy += e.n[0]->vel.x * vec[0] + e.n[1]->vel.x * vec[1] + e.n[2]->vel.x * vec[2] +
                   e.n[3]->vel.x * vec[3] + e.n[0]->vel.y * vec[4] + e.n[1]->vel.y * vec[5] +
                   e.n[2]->vel.y * vec[6] + e.n[3]->vel.y * vec[7] + e.n[0]->vel.z * vec[8] +
                   e.n[1]->vel.z * vec[9] + e.n[2]->vel.z * vec[10] + e.n[3]->vel.z * vec[11];
    }
    return kenergy * 0.5;
}

void ref_blob_destroy(ref_blob *b) {
    if (!b) return;
    delete b->solver;
    delete b->msolver;
    delete[] b->elem;
    delete[] b->node;
    delete[] b->force;
    delete[] b->pinned;
    delete[] b->rhs;
    delete[] b->params_storage;
    delete b;
}

void ref_blob_set_positions(ref_blob *b, const double *pos) {
    for (int i = 0; i < b->N; i++) {
        b->node[i].pos.x = pos[3 * i];
        b->node[i].pos.y = pos[3 * i + 1];
        b->node[i].pos.z = pos[3 * i + 2];
    }
}

static void zero_forces(ref_blob *b) {
    for (int e = 0; e < b->E; e++) b->elem[e].zero_force();
    for (int i = 0; i < b->N; i++) b->force[i].x = b->force[i].y = b->force[i].z = 0;
}

/* Blob::update_internal_forces (src/Blob.cpp:660-724); OpenMP over elements with one
 * stream per thread in the FFEA_PARALLEL_WITHIN_BLOB build, exactly one thread otherwise. */
int ref_blob_update_internal_forces(ref_blob *b, void *rngp) {
    RngStream *rng = static_cast<RngStream *>(rngp);
    int inversions = 0;
#ifdef FFEA_PARALLEL_WITHIN_BLOB
#pragma omp parallel reduction(+ : inversions)
    {
        const int tid = omp_get_thread_num();
#pragma omp for schedule(guided)
#else
    {
        const int tid = 0;
#endif
    for (int n = 0; n < b->E; n++) {
        matrix3 J, stress;
        vector12 du;
        b->elem[n].calculate_jacobian(J);
        if (b->elem[n].calc_shape_function_derivatives_and_volume(J) == FFEA_ERROR) inversions++;
        b->elem[n].create_viscosity_matrix();
        mat3_set_zero(stress);
        b->elem[n].add_shear_elastic_stress(J, stress);
        b->elem[n].add_bulk_elastic_stress(stress);
        if (b->params->calc_noise == 1) b->elem[n].add_fluctuating_stress(b->params, rng, stress, tid);
        b->elem[n].internal_stress_mag = sqrt(mat3_double_contraction_symmetric(stress));
        if (b->linear_solver != FFEA_NOMASS_CG_SOLVER) {
            b->elem[n].get_element_velocity_vector(du);
            mat12_apply(b->elem[n].viscosity_matrix, du);
        } else {
            vec12_set_zero(du);
        }
        b->elem[n].apply_stress_tensor(stress, du);
        b->elem[n].add_element_force_vector(du);
    }
    }
    b->last_inversions = inversions;
    return inversions;
}

/* Blob::aggregate_forces_and_solve up to the solve (src/Blob.cpp:3826-3923). */
static void aggregate(ref_blob *b, RngStream *rng) {
#ifdef FFEA_PARALLEL_WITHIN_BLOB
#pragma omp parallel for schedule(guided)
#endif
    for (int n = 0; n < b->N; n++)
        for (int m = 0; m < b->node[n].num_element_contributors; m++) {
            b->force[n][0] += b->node[n].force_contributions[m]->x;
            b->force[n][1] += b->node[n].force_contributions[m]->y;
            b->force[n][2] += b->node[n].force_contributions[m]->z;
        }
    if (b->params->calc_stokes == 1 && b->linear_solver != FFEA_NOMASS_CG_SOLVER) {
        /* mass-ful branch (src/Blob.cpp:3854-3877) */
#ifdef FFEA_PARALLEL_WITHIN_BLOB
#pragma omp parallel
#endif
        {
#ifdef FFEA_PARALLEL_WITHIN_BLOB
            const int thread_id = omp_get_thread_num();
#pragma omp for schedule(guided)
#else
            const int thread_id = 0;
#endif
            for (int i = 0; i < b->N; i++) {
                b->force[i].x -= b->node[i].vel.x * b->node[i].stokes_drag;
                b->force[i].y -= b->node[i].vel.y * b->node[i].stokes_drag;
                b->force[i].z -= b->node[i].vel.z * b->node[i].stokes_drag;
                if (b->params->calc_noise == 1) {
                    b->force[i].x -= RAND(-.5, .5) * sqrt((24 * b->params->kT * b->node[i].stokes_drag) / (b->params->dt));
                    b->force[i].y -= RAND(-.5, .5) * sqrt((24 * b->params->kT * b->node[i].stokes_drag) / (b->params->dt));
                    b->force[i].z -= RAND(-.5, .5) * sqrt((24 * b->params->kT * b->node[i].stokes_drag) / (b->params->dt));
                }
            }
        }
    } else if (b->params->calc_stokes == 1 && b->params->calc_noise == 1)
#ifdef FFEA_PARALLEL_WITHIN_BLOB
#pragma omp parallel
#endif
    {
#ifdef FFEA_PARALLEL_WITHIN_BLOB
        const int thread_id = omp_get_thread_num();
#pragma omp for schedule(guided)
#else
        const int thread_id = 0;
#endif
        for (int i = 0; i < b->N; i++) {
            b->force[i].x -= RAND(-.5, .5) * sqrt((24 * b->params->kT * b->node[i].stokes_drag) / (b->params->dt));
            b->force[i].y -= RAND(-.5, .5) * sqrt((24 * b->params->kT * b->node[i].stokes_drag) / (b->params->dt));
            b->force[i].z -= RAND(-.5, .5) * sqrt((24 * b->params->kT * b->node[i].stokes_drag) / (b->params->dt));
        }
    }
    /* Blob::linearise_force (src/Blob.cpp:1066-1085) */
    int nIdx[10];
    for (int i = 0; i < b->E; i++) {
        for (int j = 0; j < 10; j++) nIdx[j] = b->elem[i].n[j]->index;
        vector3 *force = b->force;
        for (int j = 0; j < 3; j++) {
            force[nIdx[0]][j] += 0.5 * (force[nIdx[4]][j] + force[nIdx[5]][j] + force[nIdx[6]][j]);
            force[nIdx[1]][j] += 0.5 * (force[nIdx[4]][j] + force[nIdx[7]][j] + force[nIdx[8]][j]);
            force[nIdx[2]][j] += 0.5 * (force[nIdx[5]][j] + force[nIdx[7]][j] + force[nIdx[9]][j]);
            force[nIdx[3]][j] += 0.5 * (force[nIdx[6]][j] + force[nIdx[8]][j] + force[nIdx[9]][j]);
        }
        for (int j = 4; j < 10; j++) force[nIdx[j]][0] = force[nIdx[j]][1] = force[nIdx[j]][2] = 0.;
    }
    for (int n = 0; n < b->num_pinned; n++) {
        int pn = b->pinned[n];
        b->force[pn].x = b->force[pn].y = b->force[pn].z = 0;
    }
    for (int i = 0; i < b->N; i++) {
        b->rhs[3 * i] = b->force[i].x;
        b->rhs[3 * i + 1] = b->force[i].y;
        b->rhs[3 * i + 2] = b->force[i].z;
    }
}

static Solver *active_solver(ref_blob *b) {
    return b->linear_solver == FFEA_NOMASS_CG_SOLVER ? static_cast<Solver *>(b->solver) : b->msolver;
}

/* Wall time per phase of the steps run so far (seconds): element loop, force gather + Stokes
 * noise + pins, solve (NoMassCGSolver::solve: build, inverse diagonal, CG), integration +
 * linearise_element -- the split of World::run's timers (src/World.cpp:1807-1809, 2089-2104). */
static double g_phase[4] = {0, 0, 0, 0};
static double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

void ref_phase_times(double out[4], int reset) {
    for (int k = 0; k < 4; k++) {
        out[k] = g_phase[k];
        if (reset) g_phase[k] = 0.0;
    }
}

/* Seconds of one SparseMatrixFixedPattern::build of the blob's viscosity matrix (the assembly
 * part of the solve phase). */
double ref_blob_time_build(ref_blob *b) {
    const double t0 = now_s();
    b->solver->V->build();
    return now_s() - t0;
}

/* Blob::update_positions (src/Blob.cpp:737-764) + euler_integrate (:3937-3967). */
int ref_blob_update_positions(ref_blob *b, void *rngp) {
    double t0 = now_s();
    aggregate(b, static_cast<RngStream *>(rngp));
    double t1 = now_s();
    g_phase[1] += t1 - t0;
    const int st = active_solver(b)->solve(b->force);
    t0 = now_s();
    g_phase[2] += t0 - t1;
    if (st == FFEA_ERROR) return -1;
    if (b->linear_solver == FFEA_NOMASS_CG_SOLVER) {
#ifdef FFEA_PARALLEL_WITHIN_BLOB
#pragma omp parallel for schedule(static)
#endif
        for (int i = 0; i < b->N; i++) {
            b->node[i].vel[0] = b->force[i][0];
            b->node[i].vel[1] = b->force[i][1];
            b->node[i].vel[2] = b->force[i][2];
            b->node[i].pos[0] += b->force[i][0] * b->params->dt;
            b->node[i].pos[1] += b->force[i][1] * b->params->dt;
            b->node[i].pos[2] += b->force[i][2] * b->params->dt;
        }
    } else {
#ifdef FFEA_PARALLEL_WITHIN_BLOB
#pragma omp parallel for schedule(static)
#endif
        for (int i = 0; i < b->N; i++) {
            b->node[i].vel[0] += b->force[i][0] * b->params->dt;
            b->node[i].vel[1] += b->force[i][1] * b->params->dt;
            b->node[i].vel[2] += b->force[i][2] * b->params->dt;
            b->node[i].pos[0] += b->node[i].vel[0] * b->params->dt;
            b->node[i].pos[1] += b->node[i].vel[1] * b->params->dt;
            b->node[i].pos[2] += b->node[i].vel[2] * b->params->dt;
        }
    }
    for (int n = 0; n < b->E; n++) b->elem[n].linearise_element();
    g_phase[3] += now_s() - t0;
    return 0;
}

void *ref_blob_node_array(ref_blob *b) { return b->node; }
void *ref_blob_element_array(ref_blob *b) { return b->elem; }
void *ref_blob_params_ptr(ref_blob *b) { return b->params; }
void ref_blob_set_max_iterations_cg(ref_blob *b, int n) { b->params->max_iterations_cg = n; }

/* One World::run iteration (src/World.cpp:1844-2043). */
int ref_world_step(ref_blob **blobs, int nblobs, void *rng) {
    int inv = 0, err = 0;
    for (int i = 0; i < nblobs; i++) zero_forces(blobs[i]);
    const double t0 = now_s();
    for (int i = 0; i < nblobs; i++) inv += ref_blob_update_internal_forces(blobs[i], rng);
    g_phase[0] += now_s() - t0;
    if (inv > 0) return inv;
    for (int i = 0; i < nblobs; i++)
        if (ref_blob_update_positions(blobs[i], rng) != 0) err = 1;
    return err ? -1 : 0;
}

int ref_blob_forces_only(ref_blob *b, void *rng, double *out_rhs) {
    zero_forces(b);
    int inv = ref_blob_update_internal_forces(b, rng);
    aggregate(b, static_cast<RngStream *>(rng));
    memcpy(out_rhs, b->rhs, sizeof(double) * 3 * (size_t)b->N);
    return inv;
}

/* The blob's solver on a caller RHS (in/out, [3*N_t]). */
int ref_blob_solve(ref_blob *b, double *x) {
    vector3 *v = new vector3[b->N];
    for (int i = 0; i < b->N; i++) v[i].assign(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    int st = active_solver(b)->solve(v);
    for (int i = 0; i < b->N; i++) {
        x[3 * i] = v[i].x;
        x[3 * i + 1] = v[i].y;
        x[3 * i + 2] = v[i].z;
    }
    delete[] v;
    return st == FFEA_OK ? 0 : -1;
}

void ref_blob_build(ref_blob *b) { b->solver->V->build(); }
int ref_blob_nnz(ref_blob *b) { return b->solver->V->get_num_nonzero_elements(); }
int ref_blob_num_rows(ref_blob *b) { return b->solver->V->get_num_rows(); }
void ref_blob_get_csr(ref_blob *b, int *key, int *col, double *val) {
    SparseMatrixFixedPattern *V = b->solver->V;
    int nr = V->get_num_rows(), nnz = V->get_num_nonzero_elements();
    sparse_entry *en = V->get_entries();
    int *k = V->get_key();
    for (int i = 0; i <= nr; i++) key[i] = k[i];
    for (int i = 0; i < nnz; i++) {
        col[i] = en[i].column_index;
        val[i] = en[i].val;
    }
}
void ref_blob_apply_matrix(ref_blob *b, const double *in, double *out) {
    vector3 *vi = new vector3[b->N], *vo = new vector3[b->N];
    for (int i = 0; i < b->N; i++) vi[i].assign(in[3 * i], in[3 * i + 1], in[3 * i + 2]);
    b->solver->V->apply(vi, vo);
    for (int i = 0; i < b->N; i++) {
        out[3 * i] = vo[i].x;
        out[3 * i + 1] = vo[i].y;
        out[3 * i + 2] = vo[i].z;
    }
    delete[] vi;
    delete[] vo;
}

void ref_blob_get(ref_blob *b, const char *what, double *out) {
    size_t N = b->N, E = b->E;
    if (!strcmp(what, "pos")) {
        for (size_t i = 0; i < N; i++)
            for (int c = 0; c < 3; c++) out[3 * i + c] = b->node[i].pos[c];
    } else if (!strcmp(what, "vel")) {
        for (size_t i = 0; i < N; i++)
            for (int c = 0; c < 3; c++) out[3 * i + c] = b->node[i].vel[c];
    } else if (!strcmp(what, "force")) {
        for (size_t i = 0; i < N; i++)
            for (int c = 0; c < 3; c++) out[3 * i + c] = b->force[i][c];
    } else if (!strcmp(what, "rhs")) {
        memcpy(out, b->rhs, sizeof(double) * 3 * N);
    } else if (!strcmp(what, "dpsi")) {
        for (size_t e = 0; e < E; e++)
            for (int k = 0; k < 12; k++) out[12 * e + k] = b->elem[e].dpsi[k];
    } else if (!strcmp(what, "vol")) {
        for (size_t e = 0; e < E; e++) out[e] = b->elem[e].vol;
    } else if (!strcmp(what, "vol_0")) {
        for (size_t e = 0; e < E; e++) out[e] = b->elem[e].vol_0;
    } else if (!strcmp(what, "J_inv_0")) {
        for (size_t e = 0; e < E; e++)
            for (int k = 0; k < 9; k++) out[9 * e + k] = b->elem[e].J_inv_0[k / 3][k % 3];
    } else if (!strcmp(what, "F_ij")) {
        for (size_t e = 0; e < E; e++)
            for (int k = 0; k < 9; k++) out[9 * e + k] = b->elem[e].F_ij[k / 3][k % 3];
    } else if (!strcmp(what, "visc")) {
        for (size_t e = 0; e < E; e++)
            for (int k = 0; k < 144; k++) out[144 * e + k] = b->elem[e].viscosity_matrix[k / 12][k % 12];
    } else if (!strcmp(what, "node_force")) {
        for (size_t e = 0; e < E; e++)
            for (int j = 0; j < 10; j++)
                for (int c = 0; c < 3; c++) out[30 * e + 3 * j + c] = b->elem[e].node_force[j][c];
    } else if (!strcmp(what, "stress_mag")) {
        for (size_t e = 0; e < E; e++) out[e] = b->elem[e].internal_stress_mag;
    } else if (!strcmp(what, "stokes_drag")) {
        for (size_t i = 0; i < N; i++) out[i] = b->node[i].stokes_drag;
    } else if (!strcmp(what, "material")) {
        for (size_t e = 0; e < E; e++) {
            out[5 * e] = b->elem[e].rho;
            out[5 * e + 1] = b->elem[e].A;
            out[5 * e + 2] = b->elem[e].B;
            out[5 * e + 3] = b->elem[e].G;
            out[5 * e + 4] = b->elem[e].E;
        }
    }
}

/* ---- steric geometry: the reference's tet_a_tetII (src/CheckTetrahedraOverlap.cpp:208-337),
 * volumeIntersectionII (src/VolumeIntersection.cpp:951-1024) and getLocalCoordinatesForLinTet
 * (src/mat_vec_fns_II.cpp:725-733).  Face.cpp needs Blob (Eigen), so the gradient loop of
 * Face::getTetraIntersectionVolumeTotalGradientAndShapeFunctions (src/Face.cpp:473-514) and
 * the scaling and spreading of VdW_solver::do_steric_interaction (src/VdW_solver.cpp:458-487)
 * are restated over those three.  Tetrahedra come as [4][3] vertex arrays. */
static void ref_tet(const double *t, grr3 *T) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 3; j++) T[i][j] = t[3 * i + j];
}

int ref_tets_overlap(const double *a, const double *b) {
    arr3 A[4], B[4];
    ref_tet(a, A);
    ref_tet(b, B);
    return tet_a_tetII(A[0], A[1], A[2], A[3], B[0], B[1], B[2], B[3]) ? 1 : 0;
}

double ref_tet_intersection(const double *a, const double *b, int want_cm, double *cm) {
    grr3 A[4], B[4], c = {0, 0, 0};
    ref_tet(a, A);
    ref_tet(b, B);
    const double v = volumeIntersectionII<geoscalar, grr3>(A[0], A[1], A[2], A[3], B[0], B[1], B[2], B[3], want_cm != 0, c);
    if (want_cm)
        for (int k = 0; k < 3; k++) cm[k] = c[k];
    return v;
}

void ref_tet_local_coords(const double *t, const double *p, double *phi) {
    grr3 T[4], P = {p[0], p[1], p[2]};
    grr4 f;
    ref_tet(t, T);
    getLocalCoordinatesForLinTet<geoscalar, grr3, grr4>(T[0], T[1], T[2], T[3], P, f);
    for (int k = 0; k < 4; k++) phi[k] = f[k];
}

/* out: [0] the energy (steric_factor V), [1..12] the forces on A's nodes, [13..24] on B's;
 * returns 1 where they apply, 0 where V is 0, -1 where the reference throws. */
int ref_steric_forces(const double *a, const double *b, double dr, double factor, double *out) {
    grr3 A[4], B[4], cm, dVdr, tA[4], tB[4];
    grr4 phi1, phi2;
    ref_tet(a, A);
    ref_tet(b, B);
    try {
        geoscalar vol = volumeIntersectionII<geoscalar, grr3>(A[0], A[1], A[2], A[3], B[0], B[1], B[2], B[3], true, cm);
        if (vol == 0) return 0;
        getLocalCoordinatesForLinTet<geoscalar, grr3, grr4>(A[0], A[1], A[2], A[3], cm, phi1);
        getLocalCoordinatesForLinTet<geoscalar, grr3, grr4>(B[0], B[1], B[2], B[3], cm, phi2);
        for (int dir = 0; dir < 3; dir++) {
            grr3 dx;
            arr3Initialise<grr3>(dx);
            dx[dir] = dr;
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 3; j++) {
                    tA[i][j] = A[i][j] + dx[j];
                    tB[i][j] = B[i][j] + dx[j];
                }
            geoscalar vm = volumeIntersectionII<geoscalar, grr3>(A[0], A[1], A[2], A[3], tB[0], tB[1], tB[2], tB[3], false, cm);
            dVdr[dir] = (vm - vol) / dr;
            vm = volumeIntersectionII<geoscalar, grr3>(B[0], B[1], B[2], B[3], tA[0], tA[1], tA[2], tA[3], false, cm);
            dVdr[dir] -= (vm - vol) / dr;
        }
        vol *= factor;
        arr3Resize<geoscalar, grr3>(factor, dVdr);
        out[0] = vol;
        for (int j = 0; j < 4; j++) {
            grr3 f1, f2;
            arr3Resize2<geoscalar, grr3>(phi1[j], dVdr, f1);
            arr3Resize2<geoscalar, grr3>(ffea_const::mOne * phi2[j], dVdr, f2);
            for (int k = 0; k < 3; k++) {
                out[1 + 3 * j + k] = f1[k];
                out[13 + 3 * j + k] = f2[k];
            }
        }
    } catch (int) {
        return -1;
    }
    return 1;
}

}  // extern "C"
