// ffh_kernels.hip -- gfx950 kernels of FFEA's time step and the C-ABI (include/ffea_hip.h)
// that drives them.  One translation unit; the device code sits in headers by subject:
//   ffh_device.h   MRG32k3a jumps and draws, fixed-order reductions, per-blob totals
//   ffh_element.h  k_element (Jacobian, shape gradients, volume + inversion, viscous and
//                  elastic + fluctuating stress, nodal forces; src/tetra_element_linear.cpp:193-490,
//                  src/Blob.cpp:638-735) and the assembly of the viscosity matrix (k_assemble,
//                  k_assemble2: each SELL block the sum of its element blocks in ascending element
//                  order, src/SparseMatrixFixedPattern.cpp:85-93)
//   ffh_forces.h   springs, constant and surface forces (World::apply_springs,
//                  Blob::apply_ctforces)
//   ffh_solve.h    k_rhs (force gather in element order, drag and noise, pins, CG start or the
//                  lumped solve; src/Blob.cpp:3826-3923), the CG kernels (k_cg_spmv / k_cg_spmv1,
//                  k_cg_update, k_blob_totals; src/NoMassCGSolver.cpp:165-292 and
//                  src/ConjugateGradientSolver.cpp:205-349), k_integrate (src/Blob.cpp:3937-3967),
//                  energies
//   ffh_box.h      World::run's periodic box (a blob whose centroid leaves it moves back in)
//   ffh_steric.h   the steric volume exclusion between two face elements (overlap test,
//                  intersection volume, its gradient; src/VolumeIntersection.cpp,
//                  src/CheckTetrahedraOverlap.cpp)
//   ffh_ssint.h    surface-surface interactions (VdW_solver): face geometry and lookup grid,
//                  and each face's force in the reference's summation order
//   ffh_mpnn.h     the learned preconditioner
// This file holds the per-handle state (ffh_blob: buffers, launches, step graphs, tuning) and
// the C-ABI.
//
// The FP64 path rounds exactly as the reference's strict build (no FMA contraction: the
// file is compiled with -ffp-contract=off); only the order of the CG dot-product sums
// differs (tree instead of sequential), deterministically.

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/ffea_hip.h"
#include "ffh_plan.h"

#include <algorithm>
#include <random>

using ffh::Mat3u;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_OR_FAIL(expr)                                                                      \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(FFH_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kBlock = 256;
constexpr int kLaneTable = 4 * 256 * 2;  // jump[kLaneTable + 2 l + comp] = A_comp^(7 l), l < 64

#include "ffh_device.h"
#include "ffh_element.h"
#include "ffh_forces.h"
#include "ffh_solve.h"
#include "ffh_box.h"
#include "ffh_steric.h"
#include "ffh_ssint.h"

#include "ffh_mpnn.h"

// ----------------------------------------------------------------------------------------
// Device buffers
// ----------------------------------------------------------------------------------------
template <typename X>
struct DBuf {
    X *p = nullptr;
    size_t n = 0;
};

}  // namespace

struct ffh_blob {
    ffh::Plan plan;
    int precision = FFH_PRECISION_FP64;
    double dt = 0, kT = 0, eps2 = 0, stokes_visc = 0;
    int maxit = 1000, calc_noise = 0, calc_stokes = 0;
    int N = 0, E = 0;
    size_t tsize = 8;
    int device = 0;  // the HIP device current at creation
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    // device arrays
    double4 *pos = nullptr;
    int4 *enodes = nullptr;
    double *jinv0 = nullptr, *vol0 = nullptr, *A = nullptr, *B = nullptr, *G = nullptr, *Em = nullptr;
    int8_t *last_sign = nullptr;
    float *elem32 = nullptr;  // FP32 element constants (mixed precision only)
    void *eforce = nullptr, *val = nullptr;
    uint32_t *nctb = nullptr, *rng_off = nullptr;
    int64_t *nt_off = nullptr;
    uint16_t *nt_cnt = nullptr;
    int32_t *slot_col = nullptr, *slice_width = nullptr;
    int64_t *slice_off = nullptr;
    uint8_t *pinned = nullptr;
    // slice assembly (k_assemble_slice): element lists, tiled contribution codes, slot tables
    int64_t *sa_eoff = nullptr, *sa_wcoff = nullptr;
    int32_t *sa_elem = nullptr, *sa_woff = nullptr, *sa_aux = nullptr;
    uint16_t *sa_code = nullptr;
    uint8_t *sa_cnt = nullptr, *sa_pos = nullptr, *slot_flags = nullptr;
    int max_upper = 0;      // most upper work tiles of a slice
    bool codes_lds = true;  // every slice's assembly codes fit the kernel's LDS area
    double *zeta = nullptr, *dinv = nullptr;
    double *f = nullptr, *r = nullptr, *x = nullptr, *p = nullptr, *q = nullptr, *z = nullptr, *vel = nullptr;
    double *parts = nullptr, *tmp_in = nullptr, *tmp_out = nullptr;
    // pinned host staging of the Solver route's vectors (ffh_blob_solve / apply_matrix), [3N]
    double *stage_in = nullptr, *stage_out = nullptr;
    int staging() {
        if (stage_in) return FFH_OK;
        const size_t bytes = 3 * (size_t)N * sizeof(double);
        if (hipHostMalloc((void **)&stage_in, bytes, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&stage_out, bytes, hipHostMallocDefault) != hipSuccess)
            return fail(FFH_E_NOMEM, "hipHostMalloc (solver staging)");
        return FFH_OK;
    }
    size_t pstride = 0;  // partials per reduction: max(slices, element blocks)
    int32_t *slice_blob = nullptr, *blob_slice_off = nullptr;
    int nblobs = 1;
    bool use_tot = false;  // some blob has more than kTotSlicesMin slices
    double *ext = nullptr;  // external forces (ffh_blob_set_external_forces)
    double *stress_mag = nullptr;  // [E] internal_stress_mag of the last element phase, when enabled
    double *stress_buf = nullptr;
    bool ext_on = false;
    // solver (ffh_blob_set_solver): mass matrix in the SELL slots + its inverse diagonal, or the
    // lumped inverse masses; densities for the kinetic energy; midpoint velocities as given
    int solver = FFH_SOLVER_CG_NOMASS;
    double *mval = nullptr, *mdinv = nullptr, *minv = nullptr, *rho = nullptr;
    std::vector<double> vel_mid;  // [3*N_t] or empty (all zero)
    // springs, constant and surface forces (ffh_blob_set_springs / _constant_forces /
    // _surface_forces): the caller's lists, and their per-node form on the device
    ffh::ForceInput fin;
    ffh::ForcePlan fplan;
    std::vector<void *> force_allocs;
    ForceArgs fdev = {};
    FaceArgs face = {};
    int nsprings = 0, nsets = 0, nmid = 0;
    double *spring_e = nullptr;
    double *trac_rows = nullptr;  // the face groups' value rows (fdev.trac, writable)
    // surface-surface interactions (ffh_blob_set_ssint): the kernels' arguments, the radix
    // sort's scratch, a host-visible error flag and the energy's per-blob-pair sums
    bool ss_on = false;
    SsArgs ss = {};
    double steric_factor = 1.0, steric_dr = 5e-3;  // SimulationParams defaults (ffh_blob_set_steric)
    std::vector<void *> ss_allocs;
    void *ss_temp = nullptr;
    size_t ss_temp_bytes = 0;
    int ss_bits = 1;
    int *ss_err_host = nullptr, *ss_err_dev = nullptr;
    double *ss_fe = nullptr, *ss_total = nullptr;
    // the energy's pair list: per face count and offset, the pairs' energies and blob pairs
    int *ss_cnt = nullptr, *ss_off = nullptr, *ss_pb = nullptr;
    double *ss_pe = nullptr;
    size_t ss_pair_cap = 0;
    void *ss_scan_temp = nullptr;
    size_t ss_scan_bytes = 0;
    void free_ssint() {
        for (void *q : ss_allocs) (void)hipFree(q);
        ss_allocs.clear();
        for (void *q : {(void *)ss_pb, (void *)ss_pe}) (void)hipFree(q);
        ss_pb = nullptr;
        ss_pe = nullptr;
        ss_pair_cap = 0;
        if (ss_err_host) (void)hipHostFree(ss_err_host);
        ss_err_host = nullptr;
        ss_err_dev = nullptr;
        ss_on = false;
        ss = SsArgs{};
        ss_temp = nullptr;
        ss_temp_bytes = 0;
    }
    // the interaction forces of this step's positions, before the element phase as in
    // World::run (src/World.cpp:1821-1925, 2002-2004); the face groups of the force kernel read them
    int run_ssint() {
        if (!ss_on) return FFH_OK;
        SsArgs a = ss;
        a.pos = pos;
        a.out = trac_rows + 3 * (size_t)fplan.ssint_row0;
        const int P = a.P, nc = a.ncell;
        hipLaunchKernelGGL(k_ss_tick, 1, 64, 0, stream, a.ctl);
        HIP_OR_FAIL(hipGetLastError());
        int rc = launch(13, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL(k_ss_geom, grid(P, 256), 256, 0, stream, e0, e1, 0, a); });
        if (rc) return rc;
        size_t tb = ss_temp_bytes;
        HIP_OR_FAIL(hipcub::DeviceRadixSort::SortPairs(ss_temp, tb, a.cell, a.sorted_cell, a.order, a.sorted, P, 0, ss_bits, stream));
        rc = launch(13, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL(k_ss_clear, grid(nc, 256), 256, 0, stream, e0, e1, 0, a); });
        if (!rc) rc = launch(13, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL(k_ss_bounds, grid(P, 256), 256, 0, stream, e0, e1, 0, a); });
        if (a.kind == FFH_SSINT_LENNARD_JONES) {
            if (!rc) rc = launch(12, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL(k_ss_forces<kSsLJ>, P, 64, 0, stream, e0, e1, 0, a); });
        } else {  // the pair potentials with the overlap screen, then the faces it deferred
            if (!rc) rc = launch(12, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL(k_ss_forces<kSsScreen>, P, 64, 0, stream, e0, e1, 0, a); });
            if (!rc) rc = launch(12, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL(k_ss_forces<kSsFull>, P, 64, 0, stream, e0, e1, 0, a); });
        }
        return rc;
    }
    int check_ssint() {
        if (ss_on && *ss_err_host == 2)
            return fail(FFH_E_ARG, "a face element too degenerate for the steric volume (a zero normal; the reference throws)");
        if (ss_on && *ss_err_host) return fail(FFH_E_ARG, "a face centroid lies outside the lookup grid (LinkedListCube)");
        return FFH_OK;
    }
    // periodic box (ffh_blob_set_box); node_off_h: the blobs' N_t offsets
    std::vector<int> node_off_h;
    int box_mask = 0, wrap_chunks = 0;
    WrapArgs wrap = {};
    std::vector<void *> box_allocs;
    // launch shapes picked by autotune(): element / assembly block size, SpMV variant
    int elem_block = kBlock, asm_block = 512, spmv_waves = 3, asm_tiles = 2;
    bool spmv_nt = false;  // non-temporal matrix loads in the SpMV
    double *tot = nullptr;
    CgCtl *ctl = nullptr;       // [nblobs]; ctl[0].inversions counts the whole batch
    CgCtl *ctl_host = nullptr;  // pinned host mirror
    uint32_t *rng_state = nullptr;
    Mat3u *jump = nullptr;
    uint32_t *wave_state = nullptr;  // [6 per 64 elements], rewritten before every element pass
    int P_slice = 0, P_elem = 0;
    // profiling
    bool profiling = false;
    double ktime[FFH_NUM_KERNELS] = {0};
    long long kcount[FFH_NUM_KERNELS] = {0};
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
    int last_iters = 0;
    double last_res = 0;
    long long steps_done = 0;  // completed position phases
    // learned preconditioner (allocated on first use)
    int precond = FFH_PRECOND_JACOBI;
    int fallback_iters = 0;  // learned iterations before the Jacobi restart (0: plain_count)
    int plain_count = -1;    // iterations of the last plain solve
    bool mp_alloced = false, mp_ready = false, mp_edges = false;
    std::vector<int32_t> blob_nreal_h;
    int32_t *blob_nreal = nullptr;
    float *mp_w = nullptr, *mp_m1 = nullptr, *mp_m2 = nullptr, *mp_g = nullptr;
    bf16_t *mp_wT = nullptr, *mp_wN = nullptr;
    float *mp_aw = nullptr, *mp_cf = nullptr;
    bf16_t *mp_H[4] = {nullptr, nullptr, nullptr, nullptr}, *mp_G[4] = {nullptr, nullptr, nullptr, nullptr};
    float *mp_feat = nullptr, *mp_u = nullptr, *mp_tgt = nullptr, *mp_dP[2] = {nullptr, nullptr}, *mp_dX = nullptr;
    float *mp_part = nullptr, *mp_loss = nullptr;
    double *r_prev = nullptr;
    int adam_t = 0;
    // harvested (r_k, x* - x_k) snapshots of plain solves
    int hv_cap = 0, hv_count = 0;
    bool hv_on = false;
    long long hv_pairs = 0;
    float *hv_r = nullptr, *hv_x = nullptr;
    std::vector<int> blob_iters;
    std::vector<double> blob_res;
    std::vector<int> blob_status;
    int cg_chunk = 16;
    // mixed path: the true residual f - K x is formed every rr_period CG iterations (and
    // whenever the recurrence residual passes the test); FFH_RR_PERIOD overrides the default
    int rr_period = 16;

    ~ffh_blob() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : force_allocs) (void)hipFree(p);
        for (void *p : box_allocs) (void)hipFree(p);
        free_ssint();
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
        for (auto &pe : pending) {
            (void)hipEventDestroy(pe.second.first);
            (void)hipEventDestroy(pe.second.second);
        }
        for (auto ev : event_pool) (void)hipEventDestroy(ev);
        for (void *p : allocs) (void)hipFree(p);
        if (ctl_host) (void)hipHostFree(ctl_host);
        if (stage_in) (void)hipHostFree(stage_in);
        if (stage_out) (void)hipHostFree(stage_out);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <typename X>
    int alloc(X *&p, size_t n) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, (n ? n : 1) * sizeof(X));
        if (e != hipSuccess) return fail(FFH_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back(q);
        p = static_cast<X *>(q);
        return FFH_OK;
    }
    template <typename X>
    int upload(X *&p, const std::vector<X> &h) {
        int rc = alloc(p, h.size());
        if (rc) return rc;
        if (!h.empty()) HIP_OR_FAIL(hipMemcpy(p, h.data(), h.size() * sizeof(X), hipMemcpyHostToDevice));
        return FFH_OK;
    }

    hipEvent_t get_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    void collect_events(bool wait) {
        std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> keep;
        for (auto &pe : pending) {
            if (!wait && hipEventQuery(pe.second.second) != hipSuccess) {
                keep.push_back(pe);
                continue;
            }
            (void)hipEventSynchronize(pe.second.second);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, pe.second.first, pe.second.second);
            ktime[pe.first] += ms;
            kcount[pe.first] += 1;
            event_pool.push_back(pe.second.first);
            event_pool.push_back(pe.second.second);
        }
        pending.swap(keep);
    }
    // Launches through fn(start, stop); with profiling on, the two events travel in the
    // dispatch packet itself (hipExtLaunchKernelGGL), so timing adds no extra packets.
    template <typename F>
    int launch(int kid, F &&fn) {
        hipEvent_t a = nullptr, b = nullptr;
        if (profiling) {
            a = get_event();
            b = get_event();
        }
        fn(a, b);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(FFH_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
        if (profiling) {
            pending.push_back({kid, {a, b}});
            if (pending.size() > 4096) collect_events(true);
        }
        return FFH_OK;
    }

    CgArgs cg_args() const {
        CgArgs v;
        v.N = N;
        v.nslices = plan.num_slices;
        v.f = f;
        v.r = r;
        v.x = x;
        v.p = p;
        v.q = q;
        v.z = z;
        v.vel = vel;
        v.dinv = solver == FFH_SOLVER_CG ? mdinv : dinv;
        v.parts_rz = parts;
        v.parts_rr = parts + 1 * pstride;
        v.parts_ff = parts + 2 * pstride;
        v.parts_pq = parts + 3 * pstride;
        v.slice_blob = slice_blob;
        v.blob_slice_off = blob_slice_off;
        v.eps2 = eps2;
        v.maxit = maxit;
        v.ctl = ctl;
        v.parts_uu = nullptr;
        v.r_prev = nullptr;
        v.parts_zrp = parts_zrp();
        v.tot = use_tot ? tot : nullptr;
        v.parts_tr = parts_tr();
        v.rr_period = rr_period;
        return v;
    }
    double *parts_energy() const { return parts + 4 * pstride; }
    double *parts_rr() const { return parts + 1 * pstride; }
    double *parts_ff() const { return parts + 2 * pstride; }
    double *parts_pq() const { return parts + 3 * pstride; }
    double *parts_uu() const { return parts + 5 * pstride; }
    double *parts_zrp() const { return parts + 6 * pstride; }
    double *parts_den() const { return parts + 7 * pstride; }
    double *parts_loss() const { return parts + 8 * pstride; }
    double *parts_tr() const { return parts + 9 * pstride; }
    // the mixed path's CG: FP32 work vectors, FP64 iterate, true residual and stopping test
    bool mixed_cg() const { return precision == FFH_PRECISION_MIXED && !massful(); }
    SpmvArgs spmv_args() const {
        SpmvArgs m;
        m.slice_off = slice_off;
        m.slice_width = slice_width;
        m.col = slot_col;
        m.val = solver == FFH_SOLVER_CG ? (const void *)mval : (const void *)val;
        return m;
    }
    bool massful() const { return solver != FFH_SOLVER_CG_NOMASS; }
    bool lumped() const { return solver == FFH_SOLVER_MASSLUMPED; }
    ElemArgs elem_args() const {
        ElemArgs g;
        g.pos = pos;
        g.enodes = enodes;
        g.jinv0 = jinv0;
        g.vol0 = vol0;
        g.A = A;
        g.B = B;
        g.G = G;
        g.Em = Em;
        g.elem32 = elem32;
        g.last_sign = last_sign;
        g.eforce = eforce;
        g.E = E;
        g.calc_noise = calc_noise;
        g.kT = kT;
        g.dt = dt;
        g.rng_state = rng_state;
        g.jump = jump;
        g.elem_base = 0;
        g.wave_state = wave_state;
        g.ctl = ctl;
        g.vel = vel;
        g.stress_mag = stress_mag;
        return g;
    }
    static int grid(int64_t n, int b = kBlock) { return (int)((n + b - 1) / b); }

    int run_element() {
        ElemArgs g = elem_args();
        const int gr = xcd_grid(grid(E, elem_block));
        if (calc_noise) {
            const int nw = (E + 63) / 64;
            hipLaunchKernelGGL(k_elem_wave_states, grid(nw), kBlock, 0, stream, (const uint32_t *)rng_state,
                               (const Mat3u *)jump, g.elem_base, nw, wave_state);
            HIP_OR_FAIL(hipGetLastError());
        }
        return launch(0, [&](hipEvent_t e0, hipEvent_t e1) {
            if (precision == FFH_PRECISION_FP64) {
                if (massful())
                    hipExtLaunchKernelGGL((k_element<double, double, true>), gr, elem_block, 0, stream, e0, e1, 0, g);
                else
                    hipExtLaunchKernelGGL((k_element<double, double, false>), gr, elem_block, 0, stream, e0, e1, 0, g);
            } else {
                if (massful())
                    hipExtLaunchKernelGGL((k_element<float, float, true>), gr, elem_block, 0, stream, e0, e1, 0, g);
                else
                    hipExtLaunchKernelGGL((k_element<float, float, false>), gr, elem_block, 0, stream, e0, e1, 0, g);
            }
        });
    }
    int run_assemble() {
        mp_edges = false;
        AsmArgs a;
        a.pos = pos;
        a.enodes = enodes;
        a.A = A;
        a.B = B;
        a.eoff = sa_eoff;
        a.elem = sa_elem;
        a.woff = sa_woff;
        a.wcoff = sa_wcoff;
        a.code = sa_code;
        a.cnt = sa_cnt;
        a.aux = sa_aux;
        a.wpos = sa_pos;
        a.slice_off = slice_off;
        a.flags = slot_flags;
        a.zeta = zeta;
        a.val = val;
        a.dinv = dinv;
        a.nslices = plan.num_slices;
        return launch(1, [&](hipEvent_t e0, hipEvent_t e1) {
            if (precision == FFH_PRECISION_FP64)
                launch_assemble<double>(a, e0, e1);
            else
                launch_assemble<float>(a, e0, e1);
        });
    }
    // asm_block threads (4, 8 or 16 waves) per slice, asm_tiles upper tiles per wave and pass;
    // the slices' codes staged in LDS when the largest fits
    template <typename T, int kWaves, bool kLds>
    void launch_assemble_w(const AsmArgs &a, hipEvent_t e0, hipEvent_t e1) {
        const int gr = P_slice;
        if (asm_tiles == 1)
            hipExtLaunchKernelGGL((k_assemble_slice<T, kWaves, 1, kLds>), gr, 64 * kWaves, 0, stream, e0, e1, 0, a);
        else if (asm_tiles == 2 || kWaves == 16)  // 16 waves: at most 2 (30 upper tiles per pass)
            hipExtLaunchKernelGGL((k_assemble_slice<T, kWaves, 2, kLds>), gr, 64 * kWaves, 0, stream, e0, e1, 0, a);
        else if constexpr (kWaves < 16)
            hipExtLaunchKernelGGL((k_assemble_slice<T, kWaves, 4, kLds>), gr, 64 * kWaves, 0, stream, e0, e1, 0, a);
    }
    template <typename T, bool kLds>
    void launch_assemble_l(const AsmArgs &a, hipEvent_t e0, hipEvent_t e1) {
        if (asm_block == 1024)
            launch_assemble_w<T, 16, kLds>(a, e0, e1);
        else if (asm_block == 512)
            launch_assemble_w<T, 8, kLds>(a, e0, e1);
        else
            launch_assemble_w<T, 4, kLds>(a, e0, e1);
    }
    template <typename T>
    void launch_assemble(const AsmArgs &a, hipEvent_t e0, hipEvent_t e1) {
        if (codes_lds)
            launch_assemble_l<T, true>(a, e0, e1);
        else
            launch_assemble_l<T, false>(a, e0, e1);
    }
    RhsArgs rhs_args(bool noise, const double *f_in) const {
        RhsArgs g;
        g.eforce = eforce;
        g.E = E;
        g.f_in = f_in;
        g.toff = nt_off;
        g.cnt = nt_cnt;
        g.ctb = nctb;
        g.zeta = zeta;
        g.pinned = pinned;
        g.rng_off = rng_off;
        g.rng_state = rng_state;
        g.jump = jump;
        g.node_base = (uint32_t)plan.elem_draws;
        g.noise = noise ? 1 : 0;
        g.kT = kT;
        g.dt = dt;
        g.ext = (ext_on && !f_in) ? ext : nullptr;
        g.vel = (massful() && calc_stokes == 1 && !f_in) ? vel : nullptr;
        g.lumped = lumped() ? minv : nullptr;
        g.fa = fdev;
        g.fa.pos = pos;
        return g;
    }
    int run_rhs(const double *f_in = nullptr) {
        if (nsets > 0 && !f_in) {  // this step's surface tractions, from the current positions
            FaceArgs fa = face;
            fa.pos = pos;
            int rc = launch(2, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL(k_face_tractions, nsets, 256, 0, stream, e0, e1, 0, fa); });
            if (rc) return rc;
        }
        if (nmid > 0 && !f_in) {  // this step's forces on the midpoints that carry one
            ForceArgs fa = fdev;
            fa.pos = pos;
            const int M = nmid;
            int rc = launch(2, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL(k_mid_forces, (M + 255) / 256, 256, 0, stream, e0, e1, 0, fa, M); });
            if (rc) return rc;
        }
        RhsArgs g = rhs_args(calc_noise == 1 && calc_stokes == 1, f_in);
        CgArgs v = cg_args();
        if (learned_solve()) v.parts_uu = parts_uu();
        int rc;
        if (precision == FFH_PRECISION_FP64)
            rc = launch(2, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL((k_rhs<double, double>), P_slice, 192, 0, stream, e0, e1, 0, g, v); });
        else if (mixed_cg())
            rc = launch(2, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL((k_rhs<float, float>), P_slice, 192, 0, stream, e0, e1, 0, g, v); });
        else
            rc = launch(2, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL((k_rhs<float, double>), P_slice, 192, 0, stream, e0, e1, 0, g, v); });
        if (rc || lumped()) return rc;
        return totals((1 << kTotFF) | (1 << kTotRZ) | (v.parts_uu ? 1 << kTotUU : 0));
    }
    int launch_spmv(int it, int mode) {
        CgArgs v = cg_args();
        SpmvArgs m = spmv_args();
        int rc = launch(3, [&](hipEvent_t e0, hipEvent_t e1) {
            const int sel = (spmv_waves == 1 ? 1 : 0) | (spmv_nt ? 2 : 0) | (precision == FFH_PRECISION_FP64 ? 4 : 0) |
                            (solver == FFH_SOLVER_CG ? 8 : 0);
            switch (sel) {
            case 0: hipExtLaunchKernelGGL((k_cg_spmv<float, false, false>), P_slice, 192, 0, stream, e0, e1, 0, m, v, it, mode); break;
            case 1: hipExtLaunchKernelGGL((k_cg_spmv1<float, false, false>), P_slice, 64, 0, stream, e0, e1, 0, m, v, it, mode); break;
            case 2: hipExtLaunchKernelGGL((k_cg_spmv<float, true, false>), P_slice, 192, 0, stream, e0, e1, 0, m, v, it, mode); break;
            case 3: hipExtLaunchKernelGGL((k_cg_spmv1<float, true, false>), P_slice, 64, 0, stream, e0, e1, 0, m, v, it, mode); break;
            case 4: hipExtLaunchKernelGGL((k_cg_spmv<double, false, false>), P_slice, 192, 0, stream, e0, e1, 0, m, v, it, mode); break;
            case 5: hipExtLaunchKernelGGL((k_cg_spmv1<double, false, false>), P_slice, 64, 0, stream, e0, e1, 0, m, v, it, mode); break;
            case 6: hipExtLaunchKernelGGL((k_cg_spmv<double, true, false>), P_slice, 192, 0, stream, e0, e1, 0, m, v, it, mode); break;
            case 7: hipExtLaunchKernelGGL((k_cg_spmv1<double, true, false>), P_slice, 64, 0, stream, e0, e1, 0, m, v, it, mode); break;
            // the mass matrix is FP64 in both precision modes
            case 8: case 12: hipExtLaunchKernelGGL((k_cg_spmv<double, false, true>), P_slice, 192, 0, stream, e0, e1, 0, m, v, it, mode); break;
            case 9: case 13: hipExtLaunchKernelGGL((k_cg_spmv1<double, false, true>), P_slice, 64, 0, stream, e0, e1, 0, m, v, it, mode); break;
            case 10: case 14: hipExtLaunchKernelGGL((k_cg_spmv<double, true, true>), P_slice, 192, 0, stream, e0, e1, 0, m, v, it, mode); break;
            default: hipExtLaunchKernelGGL((k_cg_spmv1<double, true, true>), P_slice, 64, 0, stream, e0, e1, 0, m, v, it, mode); break;
            }
        });
        if (rc) return rc;
        return totals((1 << kTotPQ) | (mixed_cg() ? 1 << kTotTR : 0));
    }
    int launch_update(int it, bool learned_next) {
        CgArgs v = cg_args();
        if (learned_next) {
            v.parts_uu = parts_uu();
            v.r_prev = r_prev;
        }
        int rc = launch(4, [&](hipEvent_t e0, hipEvent_t e1) {
            if (mixed_cg())
                hipExtLaunchKernelGGL((k_cg_update<float, true>), P_slice, 192, 0, stream, e0, e1, 0, v, it);
            else
                hipExtLaunchKernelGGL((k_cg_update<double, false>), P_slice, 192, 0, stream, e0, e1, 0, v, it);
        });
        if (rc) return rc;
        return totals((1 << kTotRR) | (1 << kTotRZ) | (learned_next ? 1 << kTotUU : 0));
    }
    int totals(int mask) {
        if (!use_tot) return FFH_OK;
        TotArgs t;
        double *pp[kTotSlots] = {parts_ff(), parts, parts_rr(), parts_pq(), parts_uu(), parts_zrp(), parts_den(), parts_tr()};
        for (int k = 0; k < kTotSlots; k++) t.parts[k] = pp[k];
        t.blob_slice_off = blob_slice_off;
        t.tot = tot;
        t.mask = mask;
        return launch(11, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL(k_blob_totals, nblobs, 1024, 0, stream, e0, e1, 0, t); });
    }
    // ---- learned preconditioner ----
    bool learned_solve() const {
        return precond == FFH_PRECOND_LEARNED && mp_ready && plain_count >= 0 && !massful();
    }
    int mp_alloc() {
        if (mp_alloced) return FFH_OK;
        const size_t n = (size_t)N;
        int rc = 0;
        blob_nreal_h.assign(nblobs, 0);
        for (int b = 0; b < nblobs; b++)
            for (int i = plan.blob_slice_off[b] * 64; i < plan.blob_slice_off[b + 1] * 64; i++)
                blob_nreal_h[b] += plan.orig_of[i] >= 0;
        rc |= upload(blob_nreal, blob_nreal_h);
        rc |= alloc(mp_w, kNumParams);
        rc |= alloc(mp_m1, kNumParams);
        rc |= alloc(mp_m2, kNumParams);
        rc |= alloc(mp_g, kNumParams);
        rc |= alloc(mp_wT, (size_t)kNLay * 64 * kHid);
        rc |= alloc(mp_wN, (size_t)kNLay * 64 * kHid);
        rc |= alloc(mp_aw, (size_t)plan.num_slots);
        rc |= alloc(mp_cf, n);
        for (int l = 0; l < 4; l++) rc |= alloc(mp_H[l], n * kHid);
        for (int l = 1; l < 4; l++) rc |= alloc(mp_G[l], n * kHid);
        rc |= alloc(mp_feat, n * kNIn);
        rc |= alloc(mp_u, 3 * n);
        rc |= alloc(mp_tgt, 3 * n);
        rc |= alloc(mp_dP[0], n * kHid);
        rc |= alloc(mp_dP[1], n * kHid);
        rc |= alloc(mp_dX, n * 64);
        const size_t nw = (n + kWgradRows - 1) / kWgradRows, ne = (n + 255) / 256;
        rc |= alloc(mp_part, std::max({(size_t)plan.num_slices * kOutGrad, nw * kLayGrad, ne * kEmbGrad}));
        rc |= alloc(mp_loss, 4096);
        rc |= alloc(r_prev, 3 * n);
        if (rc) return rc;
        mp_alloced = true;
        return FFH_OK;
    }
    MpnnArgs mp_args(bool skip_done) const {
        MpnnArgs a;
        a.N = N;
        a.nslices = plan.num_slices;
        a.slice_blob = slice_blob;
        a.blob_slice_off = blob_slice_off;
        a.blob_nreal = blob_nreal;
        a.pinned = pinned;
        a.slice_off = slice_off;
        a.slice_width = slice_width;
        a.col = slot_col;
        a.aw = mp_aw;
        a.cf = mp_cf;
        a.wf = mp_w;
        a.wT = mp_wT;
        a.wN = mp_wN;
        a.ctl = skip_done ? ctl : nullptr;
        a.tot = use_tot ? tot : nullptr;
        return a;
    }
    int mp_edges_now() {
        if (mp_edges) return FFH_OK;
        SpmvArgs m = spmv_args();
        const int g = (plan.num_slices + 3) / 4;
        const int S = plan.num_slices;
        int rc;
        if (precision == FFH_PRECISION_FP64)
            rc = launch(6, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL((k_mpnn_edges<double>), g, 256, 0, stream, e0, e1, 0, m, S, (const double *)dinv, (const double *)zeta, mp_aw, mp_cf); });
        else
            rc = launch(6, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL((k_mpnn_edges<float>), g, 256, 0, stream, e0, e1, 0, m, S, (const double *)dinv, (const double *)zeta, mp_aw, mp_cf); });
        if (rc) return rc;
        mp_edges = true;
        return FFH_OK;
    }
    int mp_pack() {
        hipLaunchKernelGGL(k_mpnn_pack, (kNLay * 64 * kHid + 255) / 256, 256, 0, stream, mp_w, mp_wT, mp_wN);
        HIP_OR_FAIL(hipGetLastError());
        return FFH_OK;
    }
    // Forward pass to z (CG).  Input: z = u = D^-1 r and the partials of |u|^2.
    // Forward pass to z.  V: the CG vectors' type (float on the mixed path); the host-facing
    // ffh_mpnn_apply works in double.
    template <typename V>
    int mpnn_apply(const V *rvec, const V *rprev, bool skip_done) {
        int rc;
        if ((rc = mp_edges_now())) return rc;
        MpnnArgs a = mp_args(skip_done);
        const int S = plan.num_slices;
        const double *uu = parts_uu();
        rc = launch(7, [&](hipEvent_t e0, hipEvent_t e1) {
            hipExtLaunchKernelGGL((k_mpnn_embed<V>), P_slice, 256, 0, stream, e0, e1, 0, a, (const V *)z, uu,
                                  (float *)nullptr, mp_H[0], (const float *)nullptr, (float *)nullptr, (double *)nullptr);
        });
        if (rc) return rc;
        bf16_t *h = mp_H[0], *hn = mp_H[1];
        for (int l = 0; l < kNLay; l++) {
            rc = launch(9, [&](hipEvent_t e0, hipEvent_t e1) {  // message + dense, fused
                hipExtLaunchKernelGGL(k_mpnn_layer, P_slice, 256, 0, stream, e0, e1, 0, a, l, (const bf16_t *)h, hn);
            });
            if (rc) return rc;
            std::swap(h, hn);
        }
        (void)S;
        double *zrp = rprev ? parts_zrp() : (double *)nullptr;
        rc = launch(10, [&](hipEvent_t e0, hipEvent_t e1) {
            hipExtLaunchKernelGGL((k_mpnn_out_cg1<V>), P_slice, 64, 0, stream, e0, e1, 0, a, (const bf16_t *)h, uu, (V *)z, rvec,
                                  rprev, parts, zrp);
        });
        if (rc) return rc;
        return totals((1 << kTotRZ) | (rprev ? 1 << kTotZRP : 0));
    }
    // One training pass over a snapshot (residual rs, target es = x* - x_k, float [3N]):
    // forward, loss into *loss_slot (device), gradient into mp_g.
    int mpnn_forward_backward(const float *rs, const float *es, float *loss_slot) {
        int rc;
        if ((rc = mp_edges_now())) return rc;
        MpnnArgs a = mp_args(false);
        const int S = plan.num_slices;
        hipLaunchKernelGGL(k_mpnn_prep, S, 192, 0, stream, S, rs, (const double *)dinv, mp_u, parts_uu());
        if ((rc = totals(1 << kTotUU))) return rc;
        hipLaunchKernelGGL((k_mpnn_embed<float>), P_slice, 256, 0, stream, a, (const float *)mp_u,
                           (const double *)parts_uu(), mp_feat, mp_H[0], es, mp_tgt, parts_den());
        if ((rc = totals(1 << kTotDEN))) return rc;
        for (int l = 1; l <= kNLay; l++) {
            hipLaunchKernelGGL((k_mpnn_spmm<bf16_t, false>), P_slice, 256, 0, stream, a, (const bf16_t *)mp_H[l - 1],
                               kHid, 0, (void *)mp_G[l], (const bf16_t *)nullptr);
            hipLaunchKernelGGL(k_mpnn_dense, S, 128, 0, stream, a, l - 1, (const bf16_t *)mp_H[l - 1],
                               (const bf16_t *)mp_G[l], mp_H[l]);
        }
        int cur = 0;
        hipLaunchKernelGGL(k_mpnn_out_train, S, 64, 0, stream, a, (const bf16_t *)mp_H[kNLay], (const float *)mp_tgt,
                           (const double *)parts_den(), 1.0f / (float)nblobs, mp_dP[cur], mp_part, parts_loss());
        hipLaunchKernelGGL(k_mpnn_sum_loss, 1, 256, 0, stream, (const double *)parts_loss(), S, loss_slot);
        hipLaunchKernelGGL(k_mpnn_reduce, 1, 256, 0, stream, (const float *)mp_part, S, kOutGrad, 1, 0, mp_g);
        const int nw = (N + kWgradRows - 1) / kWgradRows, ne = (N + 255) / 256;
        for (int l = kNLay; l >= 1; l--) {
            hipLaunchKernelGGL(k_mpnn_wgrad, nw, 128, 0, stream, N, (const bf16_t *)mp_H[l - 1], (const bf16_t *)mp_G[l],
                               (const float *)mp_dP[cur], mp_part);
            hipLaunchKernelGGL(k_mpnn_reduce, (kLayGrad + 255) / 256, 256, 0, stream, (const float *)mp_part, nw,
                               kLayGrad, 0, l - 1, mp_g);
            hipLaunchKernelGGL(k_mpnn_dgrad, S, 128, 0, stream, a, l - 1, (const float *)mp_dP[cur], mp_dX);
            hipLaunchKernelGGL((k_mpnn_spmm<float, true>), P_slice, 256, 0, stream, a, (const float *)mp_dX, 64, 32,
                               (void *)mp_dP[1 - cur], (const bf16_t *)mp_H[l - 1]);
            cur = 1 - cur;
        }
        hipLaunchKernelGGL(k_mpnn_embed_grad, ne, 256, 0, stream, N, (const float *)mp_dP[cur], (const float *)mp_feat,
                           mp_part);
        hipLaunchKernelGGL(k_mpnn_reduce, 1, 256, 0, stream, (const float *)mp_part, ne, kEmbGrad, 2, 0, mp_g);
        HIP_OR_FAIL(hipGetLastError());
        return FFH_OK;
    }
    int mpnn_adam(float lr) {
        adam_t++;
        const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)adam_t));
        const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)adam_t));
        hipLaunchKernelGGL(k_mpnn_adam, (kNumParams + 255) / 256, 256, 0, stream, mp_w, mp_m1, mp_m2,
                           (const float *)mp_g, lr, c1, c2);
        HIP_OR_FAIL(hipGetLastError());
        return mp_pack();
    }
    int harvest_snap(int k) {
        if (!hv_on || k >= hv_cap) return FFH_OK;
        const size_t n3 = 3 * (size_t)N;
        if (mixed_cg())
            hipLaunchKernelGGL((k_mpnn_harvest<float>), (int)((n3 + 255) / 256), 256, 0, stream, (int)n3, (const float *)r,
                               (const double *)x, hv_r + (size_t)k * n3, hv_x + (size_t)k * n3);
        else
            hipLaunchKernelGGL((k_mpnn_harvest<double>), (int)((n3 + 255) / 256), 256, 0, stream, (int)n3, (const double *)r,
                           (const double *)x, hv_r + (size_t)k * n3, hv_x + (size_t)k * n3);
        HIP_OR_FAIL(hipGetLastError());
        return FFH_OK;
    }
    // NoMassCGSolver::solve's loop: spmv(0), then [update(it), spmv(it+1)] until a spmv kernel
    // finds convergence (or it == maxit).  Every blob iterates on its own scalars and stops on
    // its own test; the launches go on until all have stopped.  The control blocks are polled
    // once per chunk.
    //
    // With the learned preconditioner (flexible CG), the first F iterations use it (F = the
    // fallback count, by default the iteration count of the last plain solve); iteration F + 1
    // restarts Jacobi-PCG from the current iterate (beta = 0), which then runs to the same
    // stopping test.  Snapshots for training are taken from plain solves only.
    // On the mixed path the stopping decision of iteration `it` is made in update(it) (the true
    // residual formed by spmv(it)), so there an iteration is spmv(it) followed at once by
    // update(it), step graphs included; `updated` records that update(it) is already issued.
    struct CgRun {
        bool learned, harvesting, updated;
        int F, snap0, it;
    };
    int mpnn_apply_cg(bool with_prev, bool skip_done) {
        if (mixed_cg())
            return mpnn_apply<float>((const float *)r, with_prev ? (const float *)r_prev : nullptr, skip_done);
        return mpnn_apply<double>(r, with_prev ? r_prev : nullptr, skip_done);
    }
    CgRun cg_setup() {
        CgRun c;
        c.learned = learned_solve();
        c.F = c.learned ? std::max(1, fallback_iters > 0 ? fallback_iters : plain_count) : 0;
        c.snap0 = hv_count;
        c.harvesting = hv_on && !c.learned && !massful();
        c.it = 0;
        c.updated = false;
        return c;
    }
    int cg_begin(CgRun &c) {
        int rc;
        if (c.learned && (rc = mpnn_apply_cg(false, false))) return rc;
        if ((rc = launch_spmv(0, 0))) return rc;
        if (c.harvesting && (rc = harvest_snap(c.snap0))) return rc;
        if (mixed_cg()) {
            if ((rc = launch_update(0, c.learned && 0 < c.F))) return rc;
            c.updated = true;
        }
        return FFH_OK;
    }
    int cg_iterations(CgRun &c, int count) {
        int rc;
        for (int k = 0; k < count && c.it < maxit; k++) {
            const bool lz = c.learned && c.it < c.F;
            if (!c.updated && (rc = launch_update(c.it, lz))) return rc;
            c.updated = false;
            c.it++;
            if (c.harvesting && (rc = harvest_snap(c.snap0 + c.it))) return rc;
            if (lz && (rc = mpnn_apply_cg(true, true))) return rc;
            if ((rc = launch_spmv(c.it, lz ? 1 : (c.learned && c.it == c.F + 1 ? 2 : 0)))) return rc;
            if (mixed_cg()) {
                if ((rc = launch_update(c.it, c.learned && c.it < c.F))) return rc;
                c.updated = true;
            }
        }
        return FFH_OK;
    }
    int cg_poll(bool &all_done) {
        HIP_OR_FAIL(hipMemcpyAsync(ctl_host, ctl, nblobs * sizeof(CgCtl), hipMemcpyDeviceToHost, stream));
        HIP_OR_FAIL(hipStreamSynchronize(stream));
        all_done = true;
        for (int b = 0; b < nblobs && all_done; b++) all_done = ctl_host[b].done != 0;
        return FFH_OK;
    }
    int chunk_of(const CgRun &c) const {
        return c.learned && c.it < c.F ? c.F - c.it : std::max(1, std::min(cg_chunk, maxit - c.it));
    }
    // Iterates and polls until every blob has stopped (or maxit), then the per-blob results.
    // `polled`: the launches up to c.it were already made and polled once (graph replay).
    int cg_run_rest(CgRun &c, bool polled, bool all_done) {
        int rc;
        if (!polled) all_done = false;
        while (!(polled && (all_done || c.it >= maxit))) {
            if ((rc = cg_iterations(c, chunk_of(c)))) return rc;
            if ((rc = cg_poll(all_done))) return rc;
            polled = true;
        }
        return cg_finish(c, all_done);
    }
    int run_cg() {
        CgRun c = cg_setup();
        int rc;
        if ((rc = cg_begin(c))) return rc;
        return cg_run_rest(c, false, false);
    }
    // MassLumpedSolver: the force kernel has already formed x = f / m.  Reads the control
    // blocks back (inversion count) as a CG poll would.
    int lumped_finish() {
        bool all_done = false;
        int rc;
        if ((rc = cg_poll(all_done))) return rc;
        last_iters = 0;
        last_res = 0;
        std::fill(blob_iters.begin(), blob_iters.end(), 0);
        std::fill(blob_res.begin(), blob_res.end(), 0.0);
        std::fill(blob_status.begin(), blob_status.end(), 0);
        return FFH_OK;
    }
    // The blob's Solver::solve after run_rhs.
    int run_solve() { return lumped() ? lumped_finish() : run_cg(); }
    int cg_finish(CgRun &c, bool all_done) {
        const bool learned = c.learned, harvesting = c.harvesting;
        const int snap0 = c.snap0;
        if (!all_done) return fail(FFH_E_HIP, "CG control block not finalised");
        int worst = 0, failed = -1;
        last_res = 0;
        for (int b = 0; b < nblobs; b++) {
            blob_iters[b] = ctl_host[b].iter;
            blob_res[b] = ctl_host[b].res;
            blob_status[b] = ctl_host[b].status;
            worst = std::max(worst, ctl_host[b].iter);
            last_res = std::max(last_res, ctl_host[b].res);
            if (ctl_host[b].status != 0 && failed < 0) failed = b;
        }
        last_iters = worst;
        if (!learned) {
            cg_chunk = std::max(1, std::min(worst, 64));
            plain_count = worst;
        }
        if (harvesting) {
            const int ns = std::min(hv_cap - snap0, worst);
            for (int k = 0; k < ns; k++) {
                hipLaunchKernelGGL(k_mpnn_harvest_finish, plan.num_slices, 192, 0, stream, plan.num_slices,
                                   (const int32_t *)slice_blob, (const CgCtl *)ctl, k, (const double *)x,
                                   hv_r + (size_t)(snap0 + k) * 3 * N, hv_x + (size_t)(snap0 + k) * 3 * N);
                HIP_OR_FAIL(hipGetLastError());
            }
            for (int b = 0; b < nblobs; b++) hv_pairs += std::min(blob_iters[b], ns);
            hv_count = snap0 + std::max(ns, 0);
        }
        if (failed >= 0)
            return fail(FFH_E_CG, "Conjugate gradient solver: could not converge after " + std::to_string(maxit) +
                                      " iterations" + (nblobs > 1 ? " (blob " + std::to_string(failed) + ")" : ""));
        return FFH_OK;
    }
    // gated: inside a step graph, after k_step_gate (the kernel returns at once when the gate is 0)
    int run_integrate(bool gated = false) {
        CgArgs v = cg_args();
        int adv = plan.draws_per_step > 0 ? 1 : 0;
        Mat3u j1 = plan.step_jump[0], j2 = plan.step_jump[1];
        const int *gt = gated ? step_gate : nullptr;
        return launch(5, [&](hipEvent_t e0, hipEvent_t e1) {
            hipExtLaunchKernelGGL(k_integrate, grid(N), kBlock, 0, stream, e0, e1, 0, v, pos, dt, rng_state, j1, j2, adv,
                                  massful() ? 1 : 0, gt);
        });
    }
    int run_step_gate() {
        hipLaunchKernelGGL(k_step_gate, 1, 64, 0, stream, (const CgCtl *)ctl, nblobs, lumped() ? 0 : 1,
                           (const int *)(ss_on ? ss_err_dev : nullptr), step_gate);
        HIP_OR_FAIL(hipGetLastError());
        return FFH_OK;
    }
    // ---- autotuning of launch shapes (results do not depend on them) ----
    // Best of `reps` separately timed runs after one warm-up run: the minimum is the least
    // disturbed by cold caches and by other work on the device.
    template <typename F>
    int time_ms(F &&fn, int reps, float &ms) {
        hipEvent_t a = nullptr, b = nullptr;
        HIP_OR_FAIL(hipEventCreate(&a));
        HIP_OR_FAIL(hipEventCreate(&b));
        int rc = fn();
        ms = 1e30f;
        for (int k = 0; k < reps && !rc; k++) {
            HIP_OR_FAIL(hipEventRecord(a, stream));
            rc = fn();
            HIP_OR_FAIL(hipEventRecord(b, stream));
            HIP_OR_FAIL(hipEventSynchronize(b));
            float t = 0;
            (void)hipEventElapsedTime(&t, a, b);
            ms = std::min(ms, t);
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return rc;
    }
    // Element and assembly block sizes (elements / blocks per workgroup) and the SpMV variant
    // (one wave per component vs one wave per slice), each the fastest of a few timed runs on
    // this blob's own data.  Tuning runs change no state the step depends on: the element and
    // assembly runs recompute what the next step recomputes; the solves use a synthetic force.
    int autotune() {
        int rc;
        const int sizes[3] = {64, 128, 256};
        float best = 1e30f, ms = 0;
        int pick = elem_block;
        for (int bs : sizes) {
            elem_block = bs;
            if ((rc = time_ms([&] { return run_element(); }, 5, ms))) return rc;
            if (ms < best) best = ms, pick = bs;
        }
        elem_block = pick;
        best = 1e30f;
        int pick_u = asm_tiles;
        pick = asm_block;
        for (int bs : {256, 512, 1024})
            for (int un : {1, 2, 4}) {
                if (un > 1 && (bs / 64 - 1) * (un / 2) >= max_upper) continue;  // a smaller shape covers the slice
                if (bs == 1024 && un == 4) continue;
                asm_block = bs;
                asm_tiles = un;
                if ((rc = time_ms([&] { return run_assemble(); }, 5, ms))) return rc;
                if (ms < best) best = ms, pick = bs, pick_u = un;
            }
        asm_block = pick;
        asm_tiles = pick_u;
        std::vector<double> ones(3 * (size_t)N, 1.0);
        HIP_OR_FAIL(hipMemcpyAsync(tmp_in, ones.data(), ones.size() * sizeof(double), hipMemcpyHostToDevice, stream));
        best = 1e30f;
        pick = spmv_waves;
        bool pick_nt = spmv_nt;
        const int waves[4] = {3, 1, 3, 1};
        for (int vi = 0; vi < 4; vi++) {
            const int wv = waves[vi];
            spmv_waves = wv;
            spmv_nt = vi >= 2;
            auto solve = [&] {
                int r1 = run_rhs(tmp_in);
                int r2 = r1 ? r1 : run_cg();
                return r2 == FFH_E_CG ? FFH_OK : r2;
            };
            if ((rc = time_ms(solve, 3, ms))) return rc;
            if (ms < best) best = ms, pick = wv, pick_nt = spmv_nt;
        }
        spmv_waves = pick;
        spmv_nt = pick_nt;
        plain_count = -1;
        last_iters = 0;
        last_res = 0;
        cg_chunk = 16;
        std::fill(blob_iters.begin(), blob_iters.end(), 0);
        std::fill(blob_res.begin(), blob_res.end(), 0.0);
        std::fill(blob_status.begin(), blob_status.end(), 0);
        HIP_OR_FAIL(hipStreamSynchronize(stream));
        return FFH_OK;
    }
    // Rebuilds the per-node force lists from `fin` and uploads them (replacing the previous ones).
    int upload_forces() {
        HIP_OR_FAIL(hipStreamSynchronize(stream));
        ffh::ForcePlan F;
        std::string msg = ffh::build_forces(plan, fin, F);
        if (!msg.empty()) return fail(FFH_E_ARG, msg);
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);  // the force kernels' arguments change
        graphs.clear();
        for (void *q : force_allocs) (void)hipFree(q);
        force_allocs.clear();
        fdev = ForceArgs{};
        face = FaceArgs{};
        spring_e = nullptr;
        trac_rows = nullptr;
        auto put = [&](auto *&dst, const auto &h) -> int {
            using X = std::remove_reference_t<decltype(h[0])>;
            void *q = nullptr;
            if (hipMalloc(&q, std::max<size_t>(1, h.size()) * sizeof(X)) != hipSuccess) return fail(FFH_E_NOMEM, "hipMalloc (force lists)");
            force_allocs.push_back(q);
            if (!h.empty() && hipMemcpy(q, h.data(), h.size() * sizeof(X), hipMemcpyHostToDevice) != hipSuccess)
                return fail(FFH_E_HIP, "force list upload");
            dst = reinterpret_cast<std::remove_reference_t<decltype(dst)>>(q);
            return FFH_OK;
        };
        int rc = 0;
        nsprings = (int)fin.spring_n0.size();
        nsets = F.set_off.empty() ? 0 : (int)F.set_off.size() - 1;
#define FFH_PUT(dst, h) \
    if ((rc = put(dst, h))) return rc
        if (!F.pre_off.empty()) {
            FFH_PUT(fdev.pre_off, F.pre_off);
            FFH_PUT(fdev.pre_item, F.pre_item);
            FFH_PUT(fdev.spring_nodes, F.spring_nodes);
            FFH_PUT(fdev.spring_kl, F.spring_kl);
            FFH_PUT(fdev.cforce, F.cforce);
            fdev.nconst = F.nconst;
            if (!F.rot.empty()) FFH_PUT(fdev.rot, F.rot);
        }
        if (nsprings > 0) {
            std::vector<double> z(nsprings, 0.0);
            FFH_PUT(spring_e, z);
        }
        if (!F.grp_off.empty()) {  // face groups: set tractions, or interaction corner forces
            FFH_PUT(fdev.grp_off, F.grp_off);
            FFH_PUT(fdev.grp_ent_off, F.grp_ent_off);
            FFH_PUT(fdev.grp_ent, F.grp_ent);
            std::vector<double> z3(3 * ((size_t)F.ssint_row0 + fin.ssint_faces.size()) + 3, 0.0);
            FFH_PUT(trac_rows, z3);
            fdev.trac = trac_rows;
        }
        if (nsets > 0) {
            FFH_PUT(face.set_off, F.set_off);
            FFH_PUT(face.ent_ends, F.ent_ends);
            FFH_PUT(face.set_force, F.set_force);
            std::vector<double> z((size_t)std::max(1, F.set_off.back()), 0.0);
            FFH_PUT(face.area, z);
            face.trac = trac_rows;
        }
        nmid = (int)(F.mid_ends.size() / 2);
        if (nmid > 0) {
            if (!F.mid_pre_off.empty()) FFH_PUT(fdev.mid_pre_off, F.mid_pre_off);
            if (!F.mid_grp_off.empty()) FFH_PUT(fdev.mid_grp_off, F.mid_grp_off);
            FFH_PUT(fdev.mid_ends, F.mid_ends);
            std::vector<double> z(3 * (size_t)nmid, 0.0);
            FFH_PUT(fdev.mf, z);
            FFH_PUT(fdev.fold_off, F.fold_off);
            FFH_PUT(fdev.fold_mid, F.fold_mid);
        }
#undef FFH_PUT
        fplan = std::move(F);
        return FFH_OK;
    }
    // World::run's periodic wrap at the start of a step (src/World.cpp:1850-1901)
    int run_wrap() {
        if (!box_mask) return FFH_OK;
        WrapArgs w = wrap;
        w.pos = pos;
        const int nc = wrap_chunks, nb = nblobs;
        int rc = launch(14, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL(k_wrap_partials, nc, 256, 0, stream, e0, e1, 0, w); });
        if (!rc) rc = launch(14, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL(k_wrap_decide, (nb + 63) / 64, 64, 0, stream, e0, e1, 0, w, nb); });
        if (!rc) rc = launch(14, [&](hipEvent_t e0, hipEvent_t e1) { hipExtLaunchKernelGGL(k_wrap_move, grid(N, 256), 256, 0, stream, e0, e1, 0, w); });
        return rc;
    }
    int reset_inversions() {  // a kernel, not a memset: it is also recorded into step graphs
        hipLaunchKernelGGL(k_reset_inversions, 1, 64, 0, stream, ctl);
        HIP_OR_FAIL(hipGetLastError());
        return FFH_OK;
    }
    // ---- one time step as a replayed hipGraph ----
    // The graph holds: inversion reset, element, assembly, force gather + CG start, and K CG
    // iterations (K = the last solve's count; blobs that stop earlier exit their kernels at
    // once).  One replay per step, one poll; a blob that needs more iterations continues with
    // ordinary launches, then the integration is launched.  Learned preconditioning, harvesting
    // and per-kernel profiling take the ordinary launch path.
    std::vector<std::pair<int, hipGraphExec_t>> graphs;  // by K
    bool use_graphs = true;
    int *step_gate = nullptr;  // device flag of k_step_gate
    bool integrated = false;   // the last step's graph has integrated it
    int graph_for(int K, hipGraphExec_t &ge) {
        for (auto &g : graphs)
            if (g.first == K) {
                ge = g.second;
                return FFH_OK;
            }
        hipGraph_t g = nullptr;
        HIP_OR_FAIL(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = run_wrap();
        if (!rc) rc = run_ssint();
        if (!rc) rc = reset_inversions();
        if (!rc) rc = run_element();
        if (!rc && !massful()) rc = run_assemble();
        if (!rc) rc = run_rhs();
        CgRun c = cg_setup();
        if (!rc && !lumped()) rc = cg_begin(c);
        if (!rc && !lumped()) rc = cg_iterations(c, K);
        if (!rc) rc = run_step_gate();
        if (!rc) rc = run_integrate(true);
        hipError_t e = hipStreamEndCapture(stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(FFH_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        hipGraphExec_t x = nullptr;
        e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(FFH_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        if (graphs.size() >= 8) {
            (void)hipGraphExecDestroy(graphs.front().second);
            graphs.erase(graphs.begin());
        }
        graphs.push_back({K, x});
        ge = x;
        return FFH_OK;
    }
    bool graph_step_ok() const {
        return use_graphs && !profiling && !learned_solve() && !hv_on;
    }
    // One step through a graph replay; returns like the ordinary step (inversion, CG status).
    int step_eager() {
        int rc;
        integrated = false;
        if ((rc = run_wrap())) return rc;
        if ((rc = run_ssint())) return rc;
        if ((rc = reset_inversions())) return rc;
        if ((rc = run_element())) return rc;
        if (!massful() && (rc = run_assemble())) return rc;  // mass-ful solvers use no viscosity matrix
        if ((rc = run_rhs())) return rc;
        return run_solve();  // synchronises: the inversion count is then final too
    }
    int step_graph() {
        const int K = lumped() ? 0 : std::max(1, std::min(cg_chunk, maxit));
        hipGraphExec_t ge = nullptr;
        int rc;
        if (graph_for(K, ge) != FFH_OK) {  // capture not available: ordinary launches from now on
            use_graphs = false;
            (void)hipGetLastError();
            return step_eager();
        }
        HIP_OR_FAIL(hipGraphLaunch(ge, stream));
        // the graph ends with the gated integration: it ran if the polled state passes the gate
        auto gate_passed = [&](bool all_done) {
            return all_done && ctl_host[0].inversions == 0 && !(ss_on && *ss_err_host);
        };
        if (lumped()) {
            rc = lumped_finish();
            integrated = rc == FFH_OK && gate_passed(true);
            return rc;
        }
        CgRun c = cg_setup();
        c.it = K;
        c.updated = mixed_cg();
        bool all_done = false;
        if ((rc = cg_poll(all_done))) return rc;
        integrated = gate_passed(all_done);
        return cg_run_rest(c, true, all_done);
    }
    int read_inversions(int *ninv) {
        HIP_OR_FAIL(hipMemcpyAsync(&ctl_host->inversions, &ctl->inversions, sizeof(int), hipMemcpyDeviceToHost, stream));
        HIP_OR_FAIL(hipStreamSynchronize(stream));
        int h = ctl_host->inversions;
        if (ninv) *ninv = h;
        if (h > 0) return fail(FFH_E_INVERSION, std::to_string(h) + " element(s) have inverted since the last step");
        return FFH_OK;
    }
};

extern "C" {

const char *ffh_last_error(void) { return g_err.c_str(); }

int ffh_version(void) { return (0 << 16) | 1; }

}  // extern "C"

// Common creation path of ffh_blob_create (one blob) and ffh_batch_create (the disjoint union
// of several blobs; blob_node_off marks their node ranges in the concatenated arrays).
static ffh_blob *create_impl(int num_nodes, int num_elements, const double *pos, const int *topo,
                             const double *material_si, const double *stokes_radius, int num_pinned,
                             const int *pinned, double dt, double kT, double epsilon2,
                             int max_iterations_cg, int calc_noise, int calc_stokes, double stokes_visc,
                             const unsigned long long *rng_state, int precision, int num_blobs,
                             const int *blob_node_off, bool material_meso = false) {
    if (!pos || !topo || !material_si || !stokes_radius || (num_pinned > 0 && !pinned) || !rng_state) {
        fail(FFH_E_ARG, "null argument");
        return nullptr;
    }
    if (precision != FFH_PRECISION_FP64 && precision != FFH_PRECISION_MIXED) {
        fail(FFH_E_ARG, "unknown precision mode");
        return nullptr;
    }
    if (max_iterations_cg < 1 || !(dt > 0) || !(epsilon2 >= 0)) {
        fail(FFH_E_ARG, "invalid solver parameters");
        return nullptr;
    }
    for (int i = 0; i < 6; i++)
        if (rng_state[i] >= (i < 3 ? ffh::kM1 : ffh::kM2)) {
            fail(FFH_E_ARG, "RngStream state out of range");
            return nullptr;
        }
    std::unique_ptr<ffh_blob> b(new ffh_blob());
    std::string msg = ffh::build_plan(b->plan, num_nodes, num_elements, pos, topo, material_si, stokes_radius,
                                      num_pinned, pinned, stokes_visc, calc_noise, calc_stokes, num_blobs,
                                      blob_node_off, material_meso);
    if (!msg.empty()) {
        fail(FFH_E_ARG, msg);
        return nullptr;
    }
    const ffh::Plan &P = b->plan;
    b->precision = precision;
    b->dt = dt;
    b->kT = kT;
    b->eps2 = epsilon2;
    b->maxit = max_iterations_cg;
    b->calc_noise = calc_noise;
    b->calc_stokes = calc_stokes;
    b->stokes_visc = stokes_visc;
    b->N = P.num_linear;
    b->E = P.num_elements;
    b->tsize = precision == FFH_PRECISION_FP64 ? 8 : 4;
    b->P_slice = xcd_grid(P.num_slices);
    b->P_elem = ffh_blob::grid(b->E);
    b->pstride = (size_t)std::max(P.num_slices, b->P_elem);
    b->nblobs = P.num_blobs;
    if (blob_node_off)
        b->node_off_h.assign(blob_node_off, blob_node_off + num_blobs + 1);
    else
        b->node_off_h = {0, num_nodes};
    b->blob_iters.assign(b->nblobs, 0);
    b->blob_res.assign(b->nblobs, 0.0);
    b->blob_status.assign(b->nblobs, 0);
    for (int k = 0; k < b->nblobs; k++)
        if (P.blob_slice_off[k + 1] - P.blob_slice_off[k] > kTotSlicesMin) b->use_tot = true;
    if (hipGetDevice(&b->device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess) {
        fail(FFH_E_HIP, "hipGetDevice / hipStreamCreate failed");
        return nullptr;
    }
    const int N = b->N, E = b->E;
    int rc = 0;
    rc |= b->alloc(b->pos, (size_t)N);
    if (!rc) {
        std::vector<double4> p4(N);
        for (int i = 0; i < N; i++) p4[i] = make_double4(P.pos[i], P.pos[(size_t)N + i], P.pos[2 * (size_t)N + i], 0.0);
        if (hipMemcpy(b->pos, p4.data(), p4.size() * sizeof(double4), hipMemcpyHostToDevice) != hipSuccess) rc = 1;
    }
    {
        std::vector<int4> en(E);
        for (int e = 0; e < E; e++)
            en[e] = make_int4(P.enodes[4 * (size_t)e], P.enodes[4 * (size_t)e + 1], P.enodes[4 * (size_t)e + 2],
                              P.enodes[4 * (size_t)e + 3]);
        rc |= b->upload(b->enodes, en);
    }
    rc |= b->upload(b->jinv0, P.jinv0);
    rc |= b->upload(b->vol0, P.vol0);
    rc |= b->upload(b->A, P.matA);
    rc |= b->upload(b->B, P.matB);
    rc |= b->upload(b->G, P.matG);
    rc |= b->upload(b->Em, P.matE);
    if (b->tsize == 4) {
        std::vector<float> c32(14 * (size_t)E);
        const std::vector<double> *cols[5] = {&P.vol0, &P.matA, &P.matB, &P.matG, &P.matE};
        for (size_t i = 0; i < 9 * (size_t)E; i++) c32[i] = (float)P.jinv0[i];
        for (int r = 0; r < 5; r++)
            for (int e = 0; e < E; e++) c32[(9 + r) * (size_t)E + e] = (float)(*cols[r])[e];
        rc |= b->upload(b->elem32, c32);
    }
    {
        std::vector<int8_t> sg(E);
        for (int e = 0; e < E; e++) sg[e] = P.last_det[e] > 0.0 ? 1 : (P.last_det[e] < 0.0 ? -1 : 0);
        rc |= b->upload(b->last_sign, sg);
    }
    {
        char *ef = nullptr, *vv = nullptr;
        rc |= b->alloc(ef, 12 * (size_t)E * b->tsize);
        rc |= b->alloc(vv, 9 * (size_t)P.num_slots * b->tsize);
        b->eforce = ef;
        b->val = vv;
        // padding slots stay zero for good; real slots are rewritten by every assembly
        if (!rc && hipMemset(vv, 0, 9 * (size_t)P.num_slots * b->tsize) != hipSuccess) rc = 1;
    }
    rc |= b->upload(b->sa_eoff, P.sa_eoff);
    rc |= b->upload(b->sa_elem, P.sa_elem);
    rc |= b->upload(b->sa_woff, P.sa_woff);
    rc |= b->upload(b->sa_wcoff, P.sa_wcoff);
    rc |= b->upload(b->sa_code, P.sa_code);
    rc |= b->upload(b->sa_cnt, P.sa_cnt);
    rc |= b->upload(b->sa_aux, P.sa_aux);
    rc |= b->upload(b->sa_pos, P.sa_pos);
    rc |= b->upload(b->slot_flags, P.slot_flags);
    for (int s = 0; s < P.num_slices; s++) {
        const int w0 = P.sa_woff[s], w1 = P.sa_woff[s + 1];
        b->max_upper = std::max(b->max_upper, w1 - w0 - 1);
        if (P.sa_wcoff[w1] - P.sa_wcoff[w0] > kAsmCodeLds) b->codes_lds = false;
    }
    rc |= b->upload(b->nt_off, P.ntile.off);
    rc |= b->upload(b->nt_cnt, P.ntile.cnt);
    rc |= b->upload(b->nctb, P.ntile.code);
    rc |= b->upload(b->rng_off, P.rng_node_off);
    rc |= b->upload(b->slot_col, P.slot_col);
    rc |= b->upload(b->slice_width, P.slice_width);
    rc |= b->upload(b->slice_off, P.slice_off);
    rc |= b->upload(b->pinned, P.pinned);
    rc |= b->upload(b->zeta, P.zeta);
    rc |= b->alloc(b->dinv, 3 * (size_t)N);
    rc |= b->alloc(b->step_gate, 1);
    rc |= b->alloc(b->f, 3 * (size_t)N);
    rc |= b->alloc(b->r, 3 * (size_t)N);
    rc |= b->alloc(b->x, 3 * (size_t)N);
    rc |= b->alloc(b->p, 3 * (size_t)N);
    rc |= b->alloc(b->z, 3 * (size_t)N);
    rc |= b->alloc(b->q, 3 * (size_t)N);
    rc |= b->alloc(b->vel, 3 * (size_t)N);
    rc |= b->alloc(b->tmp_in, 3 * (size_t)N);
    rc |= b->alloc(b->tmp_out, 3 * (size_t)N);
    rc |= b->alloc(b->parts, 10 * b->pstride);
    rc |= b->upload(b->slice_blob, P.slice_blob);
    rc |= b->upload(b->blob_slice_off, P.blob_slice_off);
    rc |= b->alloc(b->ctl, (size_t)b->nblobs);
    rc |= b->alloc(b->tot, (size_t)b->nblobs * kTotSlots);
    rc |= b->alloc(b->rng_state, 6);
    rc |= b->alloc(b->wave_state, 6 * (((size_t)E + 63) / 64) + 6);
    rc |= b->upload(b->jump, P.jump);
    if (rc) {
        if (g_err.empty()) fail(FFH_E_HIP, "device allocation or upload failed");
        return nullptr;
    }
    if (hipHostMalloc((void **)&b->ctl_host, b->nblobs * sizeof(CgCtl), hipHostMallocDefault) != hipSuccess) {
        fail(FFH_E_HIP, "hipHostMalloc failed");
        return nullptr;
    }
    std::vector<CgCtl> c(b->nblobs);
    std::memset(c.data(), 0, c.size() * sizeof(CgCtl));
    uint32_t st[6];
    for (int i = 0; i < 6; i++) st[i] = (uint32_t)rng_state[i];
    if (hipMemcpy(b->ctl, c.data(), c.size() * sizeof(CgCtl), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->rng_state, st, sizeof(st), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(b->vel, 0, 3 * (size_t)N * sizeof(double)) != hipSuccess ||
        hipMemset(b->x, 0, 3 * (size_t)N * sizeof(double)) != hipSuccess) {
        fail(FFH_E_HIP, "device initialisation failed");
        return nullptr;
    }
    // NoMassCGSolver::init builds the element matrices at the current positions; so does the
    // first assembly here (no noise draws are consumed: the element kernel's forces are unused).
    {
        int cn = b->calc_noise;
        b->calc_noise = 0;
        int r1 = b->run_element();
        b->calc_noise = cn;
        if (r1 || b->run_assemble() || hipStreamSynchronize(b->stream) != hipSuccess) {
            if (g_err.empty()) fail(FFH_E_HIP, "initial assembly failed");
            return nullptr;
        }
    }
    if (const char *rp = std::getenv("FFH_RR_PERIOD")) b->rr_period = std::max(0, std::atoi(rp));
    const char *tune = std::getenv("FFH_AUTOTUNE");
    if (!(tune && tune[0] == '0') && b->autotune() != FFH_OK) {
        if (g_err.empty()) fail(FFH_E_HIP, "autotuning failed");
        return nullptr;
    }
    return b.release();
}

extern "C" {

ffh_blob *ffh_blob_create(int num_nodes, int num_elements, const double *pos, const int *topo,
                          const double *material_si, const double *stokes_radius, int num_pinned,
                          const int *pinned, double dt, double kT, double epsilon2,
                          int max_iterations_cg, int calc_noise, int calc_stokes, double stokes_visc,
                          const unsigned long long *rng_state, int precision) {
    g_err.clear();
    return create_impl(num_nodes, num_elements, pos, topo, material_si, stokes_radius, num_pinned, pinned, dt,
                       kT, epsilon2, max_iterations_cg, calc_noise, calc_stokes, stokes_visc, rng_state, precision,
                       1, nullptr);
}

ffh_blob *ffh_blob_create_meso(int num_nodes, int num_elements, const double *pos, const int *topo,
                               const double *material_meso, const double *stokes_radius, int num_pinned,
                               const int *pinned, double dt, double kT, double epsilon2,
                               int max_iterations_cg, int calc_noise, int calc_stokes, double stokes_visc,
                               const unsigned long long *rng_state, int precision) {
    g_err.clear();
    return create_impl(num_nodes, num_elements, pos, topo, material_meso, stokes_radius, num_pinned, pinned, dt,
                       kT, epsilon2, max_iterations_cg, calc_noise, calc_stokes, stokes_visc, rng_state, precision,
                       1, nullptr, true);
}

ffh_blob *ffh_batch_create(int num_blobs, const int *num_nodes, const int *num_elements, const double *pos,
                           const int *topo, const double *material_si, const double *stokes_radius,
                           const int *num_pinned, const int *pinned, double dt, double kT, double epsilon2,
                           int max_iterations_cg, int calc_noise, int calc_stokes, double stokes_visc,
                           const unsigned long long *rng_state, int precision) {
    g_err.clear();
    if (num_blobs < 1 || !num_nodes || !num_elements || !num_pinned) {
        fail(FFH_E_ARG, "null argument or no blobs");
        return nullptr;
    }
    // concatenate: node indices of blob b's topology and pins are shifted by its node offset
    std::vector<int> noff(num_blobs + 1, 0);
    int64_t tn = 0, te = 0, tp = 0;
    for (int b = 0; b < num_blobs; b++) {
        if (num_nodes[b] <= 0 || num_elements[b] <= 0 || num_pinned[b] < 0) {
            fail(FFH_E_ARG, "blob " + std::to_string(b) + ": needs nodes and elements");
            return nullptr;
        }
        tn += num_nodes[b];
        te += num_elements[b];
        tp += num_pinned[b];
        if (tn >= (int64_t)1 << 30 || te >= (int64_t)1 << 27) {
            fail(FFH_E_ARG, "batch too large");
            return nullptr;
        }
        noff[b + 1] = (int)tn;
    }
    if (!topo || (tp > 0 && !pinned)) {
        fail(FFH_E_ARG, "null argument");
        return nullptr;
    }
    std::vector<int> gtopo(10 * (size_t)te), gpin((size_t)tp);
    int64_t e0 = 0, p0 = 0;
    for (int b = 0; b < num_blobs; b++) {
        for (int64_t k = 0; k < 10 * (int64_t)num_elements[b]; k++) {
            int v = topo[10 * e0 + k];
            if (v < 0 || v >= num_nodes[b]) {
                fail(FFH_E_ARG, "blob " + std::to_string(b) + ": topology node index out of bounds");
                return nullptr;
            }
            gtopo[10 * e0 + k] = v + noff[b];
        }
        for (int k = 0; k < num_pinned[b]; k++) {
            int v = pinned[p0 + k];
            if (v < 0 || v >= num_nodes[b]) {
                fail(FFH_E_ARG, "blob " + std::to_string(b) + ": pinned node index out of bounds");
                return nullptr;
            }
            gpin[p0 + k] = v + noff[b];
        }
        e0 += num_elements[b];
        p0 += num_pinned[b];
    }
    return create_impl((int)tn, (int)te, pos, gtopo.data(), material_si, stokes_radius, (int)tp, gpin.data(), dt,
                       kT, epsilon2, max_iterations_cg, calc_noise, calc_stokes, stokes_visc, rng_state, precision,
                       num_blobs, noff.data());
}

int ffh_batch_num_blobs(ffh_blob *b) { return b ? b->nblobs : fail(FFH_E_ARG, "null blob"); }

int ffh_batch_last_cg(ffh_blob *b, int *iterations, double *residual2, int *status) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    for (int k = 0; k < b->nblobs; k++) {
        if (iterations) iterations[k] = b->blob_iters[k];
        if (residual2) residual2[k] = b->blob_res[k];
        if (status) status[k] = b->blob_status[k] == 0 ? FFH_OK : FFH_E_CG;
    }
    return FFH_OK;
}

void ffh_blob_destroy(ffh_blob *b) { delete b; }

int ffh_blob_update_internal_forces(ffh_blob *b, int *num_inversions) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    int rc;
    if ((rc = b->run_wrap())) return rc;
    if ((rc = b->run_ssint())) return rc;
    if ((rc = b->reset_inversions())) return rc;
    if ((rc = b->run_element())) return rc;
    if (!b->massful() && (rc = b->run_assemble())) return rc;
    if ((rc = b->read_inversions(num_inversions))) return rc;
    return b->check_ssint();
}

int ffh_blob_update_positions(ffh_blob *b) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    int rc;
    if ((rc = b->run_rhs())) return rc;
    int cg = b->run_solve();
    if ((rc = b->run_integrate())) return rc;
    b->steps_done++;
    return cg;
}

int ffh_blob_step(ffh_blob *b, int nsteps) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    for (int s = 0; s < nsteps; s++) {
        int rc, cg;
        cg = b->graph_step_ok() ? b->step_graph() : b->step_eager();  // both synchronise
        if (cg != FFH_OK && cg != FFH_E_CG) return cg;
        if ((rc = b->check_ssint())) return rc;
        if (b->ctl_host->inversions > 0) {
            // the reference aborts before the position phase (src/World.cpp:2025-2028)
            int ninv = 0;
            return b->read_inversions(&ninv);
        }
        if (!b->integrated && (rc = b->run_integrate())) return rc;
        b->steps_done++;
        if (cg) return cg;
    }
    return FFH_OK;
}

int ffh_blob_tuning(ffh_blob *b, int *element_block, int *assembly_block, int *assembly_tiles,
                    int *spmv_waves_per_slice, int *spmv_nontemporal) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    if (element_block) *element_block = b->elem_block;
    if (assembly_block) *assembly_block = b->asm_block;
    if (assembly_tiles) *assembly_tiles = b->asm_tiles;
    if (spmv_waves_per_slice) *spmv_waves_per_slice = b->spmv_waves;
    if (spmv_nontemporal) *spmv_nontemporal = b->spmv_nt ? 1 : 0;
    return FFH_OK;
}

int ffh_blob_set_tuning(ffh_blob *b, int element_block, int assembly_block, int assembly_tiles,
                        int spmv_waves_per_slice, int spmv_nontemporal) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    auto ok = [](int v) { return v == 64 || v == 128 || v == 256; };
    if (!ok(element_block) || (assembly_block != 256 && assembly_block != 512 && assembly_block != 1024) ||
        (spmv_waves_per_slice != 1 && spmv_waves_per_slice != 3) ||
        (assembly_tiles != 1 && assembly_tiles != 2 && assembly_tiles != 4) ||
        (assembly_block == 1024 && assembly_tiles == 4))
        return fail(FFH_E_ARG, "the element block size must be 64, 128 or 256, the assembly workgroup 256, 512 or "
                               "1024 threads with 1, 2 or 4 upper tiles per wave (1 or 2 with 1024), and SpMV waves per slice 1 or 3");
    b->elem_block = element_block;
    b->asm_block = assembly_block;
    b->asm_tiles = assembly_tiles;
    b->spmv_waves = spmv_waves_per_slice;
    b->spmv_nt = spmv_nontemporal != 0;
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);
    b->graphs.clear();
    return FFH_OK;
}

int ffh_blob_set_external_forces(ffh_blob *b, const double *force) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);  // the force gather's arguments change
    b->graphs.clear();
    if (!force) {
        b->ext_on = false;
        return FFH_OK;
    }
    const ffh::Plan &P = b->plan;
    const size_t NT = (size_t)P.num_nodes_total;
    std::vector<double> full(force, force + 3 * NT);
    for (size_t k = 0; k < 3 * NT; k++)
        if (!std::isfinite(full[k])) return fail(FFH_E_ARG, "non-finite external force");
    // linearise_force (src/Blob.cpp:1066-1085): element by element, each vertex takes half the
    // sum of the forces on its three edge midpoints, then the element's midpoints are zeroed
    for (int e = 0; e < P.num_elements; e++) {
        int v[4];
        for (int a = 0; a < 4; a++) v[a] = P.orig_of[P.enodes[4 * (size_t)e + a]];
        const int32_t *m = &P.emid[6 * (size_t)e];
        for (int j = 0; j < 3; j++) {
            full[3 * (size_t)v[0] + j] += 0.5 * (full[3 * (size_t)m[0] + j] + full[3 * (size_t)m[1] + j] + full[3 * (size_t)m[2] + j]);
            full[3 * (size_t)v[1] + j] += 0.5 * (full[3 * (size_t)m[0] + j] + full[3 * (size_t)m[3] + j] + full[3 * (size_t)m[4] + j]);
            full[3 * (size_t)v[2] + j] += 0.5 * (full[3 * (size_t)m[1] + j] + full[3 * (size_t)m[3] + j] + full[3 * (size_t)m[5] + j]);
            full[3 * (size_t)v[3] + j] += 0.5 * (full[3 * (size_t)m[2] + j] + full[3 * (size_t)m[4] + j] + full[3 * (size_t)m[5] + j]);
        }
        for (int k = 0; k < 6; k++)
            for (int j = 0; j < 3; j++) full[3 * (size_t)m[k] + j] = 0.0;
    }
    std::vector<double> lin(3 * (size_t)b->N, 0.0);
    for (int i = 0; i < b->N; i++)
        if (P.orig_of[i] >= 0)
            for (int j = 0; j < 3; j++) lin[3 * (size_t)i + j] = full[3 * (size_t)P.orig_of[i] + j];
    if (!b->ext) {
        int rc = b->alloc(b->ext, lin.size());
        if (rc) return rc;
    }
    HIP_OR_FAIL(hipMemcpyAsync(b->ext, lin.data(), lin.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    b->ext_on = true;
    return FFH_OK;
}

int ffh_blob_set_graphs(ffh_blob *b, int enable) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    b->use_graphs = enable != 0;
    return FFH_OK;
}

int ffh_blob_steps_done(ffh_blob *b, long long *steps) {
    if (!b || !steps) return fail(FFH_E_ARG, "null argument");
    *steps = b->steps_done;
    return FFH_OK;
}

int ffh_blob_solve(ffh_blob *b, double *xio) {
    if (!b || !xio) return fail(FFH_E_ARG, "null argument");
    const ffh::Plan &P = b->plan;
    for (size_t k = 0; k < P.mid_index.size(); k++) {
        int m = P.mid_index[k];
        if (xio[3 * m] != 0.0 || xio[3 * m + 1] != 0.0 || xio[3 * m + 2] != 0.0)
            return fail(FFH_E_ARG, "midpoint rows of the force vector must be zero");
    }
    if (int rc0 = b->staging()) return rc0;
    // linear rows compacted straight into pinned memory (padding rows zero), one DMA each way
    double *fl = b->stage_in, *xl = b->stage_out;
    for (int i = 0; i < b->N; i++) {
        const int o = P.orig_of[i];
        for (int c = 0; c < 3; c++) fl[3 * (size_t)i + c] = o >= 0 ? xio[3 * (size_t)o + c] : 0.0;
    }
    const size_t bytes = 3 * (size_t)b->N * sizeof(double);
    HIP_OR_FAIL(hipMemcpyAsync(b->tmp_in, fl, bytes, hipMemcpyHostToDevice, b->stream));
    if (int rc0 = b->run_rhs(b->tmp_in)) return rc0;
    int rc = b->run_solve();
    HIP_OR_FAIL(hipMemcpyAsync(xl, b->x, bytes, hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    for (int i = 0; i < b->N; i++)
        if (P.orig_of[i] >= 0)
            for (int c = 0; c < 3; c++) xio[3 * (size_t)P.orig_of[i] + c] = xl[3 * (size_t)i + c];
    return rc;
}

int ffh_blob_apply_matrix(ffh_blob *b, const double *in, double *result) {
    if (!b || !in || !result) return fail(FFH_E_ARG, "null argument");
    if (b->massful()) return fail(FFH_E_ARG, "apply_matrix: the viscosity matrix of the massless solver only");
    const ffh::Plan &P = b->plan;
    if (int rc0 = b->staging()) return rc0;
    double *il = b->stage_in, *ol = b->stage_out;
    for (int i = 0; i < b->N; i++) {
        const int o = P.orig_of[i];
        for (int c = 0; c < 3; c++) il[3 * (size_t)i + c] = o >= 0 ? in[3 * (size_t)o + c] : 0.0;
    }
    const size_t bytes = 3 * (size_t)b->N * sizeof(double);
    HIP_OR_FAIL(hipMemcpyAsync(b->tmp_in, il, bytes, hipMemcpyHostToDevice, b->stream));
    SpmvArgs m = b->spmv_args();
    double *ti = b->tmp_in, *to = b->tmp_out;
    int N = b->N;
    if (b->precision == FFH_PRECISION_FP64)
        hipLaunchKernelGGL((k_apply<double>), ffh_blob::grid(N), kBlock, 0, b->stream, m, N, (const double *)ti, to);
    else
        hipLaunchKernelGGL((k_apply<float>), ffh_blob::grid(N), kBlock, 0, b->stream, m, N, (const double *)ti, to);
    HIP_OR_FAIL(hipGetLastError());
    HIP_OR_FAIL(hipMemcpyAsync(ol, b->tmp_out, bytes, hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    // midpoint rows (and nodes of no element) are identity; linear rows from the device
    std::memcpy(result, in, 3 * (size_t)P.num_nodes_total * sizeof(double));
    for (int i = 0; i < b->N; i++)
        if (P.orig_of[i] >= 0)
            for (int c = 0; c < 3; c++) result[3 * (size_t)P.orig_of[i] + c] = ol[3 * (size_t)i + c];
    return FFH_OK;
}

int ffh_blob_set_stress_output(ffh_blob *b, int enable) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);  // the element kernel's arguments change
    b->graphs.clear();
    if (!enable) {
        b->stress_mag = nullptr;  // the buffer stays with the handle's allocations
        return FFH_OK;
    }
    if (!b->stress_buf) {
        int rc = b->alloc(b->stress_buf, (size_t)b->E);
        if (rc) return rc;
        HIP_OR_FAIL(hipMemset(b->stress_buf, 0, (size_t)b->E * sizeof(double)));
    }
    b->stress_mag = b->stress_buf;
    return FFH_OK;
}

int ffh_blob_get_stress_magnitudes(ffh_blob *b, double *out) {
    if (!b || !out) return fail(FFH_E_ARG, "null argument");
    if (!b->stress_mag) return fail(FFH_E_ARG, "stress output is not enabled (ffh_blob_set_stress_output)");
    HIP_OR_FAIL(hipMemcpyAsync(out, b->stress_mag, (size_t)b->E * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    return FFH_OK;
}

int ffh_blob_update_matrix(ffh_blob *b) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    if (b->massful()) return fail(FFH_E_ARG, "update_matrix: the viscosity matrix of the massless solver only");
    if (int rc = b->run_assemble()) return rc;
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    return FFH_OK;
}

int ffh_blob_set_positions(ffh_blob *b, const double *pos) {
    if (!b || !pos) return fail(FFH_E_ARG, "null argument");
    const int N = b->N;
    std::vector<double4> h(N, make_double4(0.0, 0.0, 0.0, 0.0));
    for (int i = 0; i < N; i++)
        if (b->plan.orig_of[i] >= 0) {
            const double *q = pos + 3 * (size_t)b->plan.orig_of[i];
            h[i] = make_double4(q[0], q[1], q[2], 0.0);
        }
    HIP_OR_FAIL(hipMemcpyAsync(b->pos, h.data(), h.size() * sizeof(double4), hipMemcpyHostToDevice, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    return FFH_OK;
}

int ffh_blob_get_positions(ffh_blob *b, double *pos) {
    if (!b || !pos) return fail(FFH_E_ARG, "null argument");
    const int N = b->N;
    const ffh::Plan &P = b->plan;
    std::vector<double4> h(N);
    HIP_OR_FAIL(hipMemcpyAsync(h.data(), b->pos, h.size() * sizeof(double4), hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    for (int i = 0; i < N; i++)
        if (P.orig_of[i] >= 0) {
            double *q = pos + 3 * (size_t)P.orig_of[i];
            q[0] = h[i].x;
            q[1] = h[i].y;
            q[2] = h[i].z;
        }
    // linearise_element (src/tetra_element_linear.cpp:567-591)
    for (size_t k = 0; k < P.mid_index.size(); k++)
        for (int c = 0; c < 3; c++)
            pos[3 * (size_t)P.mid_index[k] + c] = .5 * (pos[3 * (size_t)P.mid_a[k] + c] + pos[3 * (size_t)P.mid_b[k] + c]);
    return FFH_OK;
}

static int get_node_vector(ffh_blob *b, const double *dev, double *out) {
    const ffh::Plan &P = b->plan;
    std::vector<double> h(3 * (size_t)b->N);
    HIP_OR_FAIL(hipMemcpyAsync(h.data(), dev, h.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    std::memset(out, 0, sizeof(double) * 3 * (size_t)P.num_nodes_total);
    for (int i = 0; i < b->N; i++)
        if (P.orig_of[i] >= 0)
            for (int c = 0; c < 3; c++) out[3 * (size_t)P.orig_of[i] + c] = h[3 * i + c];
    return FFH_OK;
}

int ffh_blob_get_velocities(ffh_blob *b, double *vel) {
    if (!b || !vel) return fail(FFH_E_ARG, "null argument");
    int rc = get_node_vector(b, b->vel, vel);
    if (rc || b->vel_mid.empty() || !b->massful()) return rc;
    for (int m : b->plan.mid_index)
        for (int c = 0; c < 3; c++) vel[3 * (size_t)m + c] = b->vel_mid[3 * (size_t)m + c];
    return FFH_OK;
}

int ffh_blob_set_velocities(ffh_blob *b, const double *vel) {
    if (!b || !vel) return fail(FFH_E_ARG, "null argument");
    const ffh::Plan &P = b->plan;
    std::vector<double> h(3 * (size_t)b->N, 0.0);
    for (int i = 0; i < b->N; i++)
        if (P.orig_of[i] >= 0)
            for (int c = 0; c < 3; c++) h[3 * (size_t)i + c] = vel[3 * (size_t)P.orig_of[i] + c];
    b->vel_mid.assign(3 * (size_t)P.num_nodes_total, 0.0);
    bool any = false;
    for (int m : P.mid_index)
        for (int c = 0; c < 3; c++) any |= (b->vel_mid[3 * (size_t)m + c] = vel[3 * (size_t)m + c]) != 0.0;
    if (!any) b->vel_mid.clear();
    HIP_OR_FAIL(hipMemcpyAsync(b->vel, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    return FFH_OK;
}

int ffh_blob_set_springs(ffh_blob *b, int num_springs, const int *node0, const int *node1, const double *k,
                         const double *l) {
    if (!b || num_springs < 0 || (num_springs > 0 && (!node0 || !node1 || !k || !l))) return fail(FFH_E_ARG, "null argument");
    for (int i = 0; i < num_springs; i++)
        if (!(k[i] >= 0) || !(l[i] >= 0) || !std::isfinite(k[i]) || !std::isfinite(l[i]))
            return fail(FFH_E_ARG, "spring " + std::to_string(i) + ": k and l must be finite and >= 0");
    ffh::ForceInput keep = b->fin;
    b->fin.spring_n0.assign(node0, node0 + num_springs);
    b->fin.spring_n1.assign(node1, node1 + num_springs);
    b->fin.spring_k.assign(k, k + num_springs);
    b->fin.spring_l.assign(l, l + num_springs);
    int rc = b->upload_forces();
    if (rc) {
        b->fin = keep;
        (void)b->upload_forces();
    }
    return rc;
}

int ffh_blob_spring_energy(ffh_blob *b, double *energy) {
    if (!b || !energy) return fail(FFH_E_ARG, "null argument");
    *energy = 0.0;
    if (b->nsprings == 0) return FFH_OK;
    ForceArgs fa = b->fdev;
    fa.pos = b->pos;
    hipLaunchKernelGGL(k_spring_energy, (b->nsprings + 255) / 256, 256, 0, b->stream, fa, b->nsprings, b->spring_e);
    HIP_OR_FAIL(hipGetLastError());
    std::vector<double> e(b->nsprings);
    HIP_OR_FAIL(hipMemcpyAsync(e.data(), b->spring_e, e.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    // springfieldenergy[blob0][blob1] in spring order, then summed over the pairs
    const int nb = b->nblobs;
    std::vector<double> pair((size_t)nb * nb, 0.0);
    for (int s = 0; s < b->nsprings; s++)
        pair[(size_t)b->fplan.spring_blob[2 * s] * nb + b->fplan.spring_blob[2 * s + 1]] += e[s];
    double tot = 0.0;
    for (double v : pair) tot += v;
    *energy = tot;
    return FFH_OK;
}

int ffh_blob_set_constant_forces(ffh_blob *b, int num_forces, const int *node, const double *force) {
    if (!b || num_forces < 0 || (num_forces > 0 && (!node || !force))) return fail(FFH_E_ARG, "null argument");
    for (int i = 0; i < 3 * num_forces; i++)
        if (!std::isfinite(force[i])) return fail(FFH_E_ARG, "non-finite constant force");
    ffh::ForceInput keep = b->fin;
    b->fin.const_node.assign(node, node + num_forces);
    b->fin.const_force.assign(force, force + 3 * (size_t)num_forces);
    int rc = b->upload_forces();
    if (rc) {
        b->fin = keep;
        (void)b->upload_forces();
    }
    return rc;
}

int ffh_blob_set_rotational_forces(ffh_blob *b, int num_forces, const int *node, const double *F, const double *axis,
                                   const int *torque) {
    if (!b || num_forces < 0 || (num_forces > 0 && (!node || !F || !axis || !torque))) return fail(FFH_E_ARG, "null argument");
    for (int i = 0; i < num_forces; i++) {
        bool ok = std::isfinite(F[i]);
        for (int k = 0; k < 6; k++) ok = ok && std::isfinite(axis[6 * i + k]);
        if (!ok) return fail(FFH_E_ARG, "non-finite rotational force or axis");
    }
    ffh::ForceInput keep = b->fin;
    b->fin.rot_node.assign(node, node + num_forces);
    b->fin.rot_F.assign(F, F + num_forces);
    b->fin.rot_axis.assign(axis, axis + 6 * (size_t)num_forces);
    b->fin.rot_torque.assign(torque, torque + num_forces);
    int rc = b->upload_forces();
    if (rc) {
        b->fin = keep;
        (void)b->upload_forces();
    }
    return rc;
}

int ffh_blob_set_ssint(ffh_blob *b, int kind, int num_faces, const int *face_nodes, const int *face_opposite,
                       const int *face_type, int num_types, const double *emin, const double *rmin, const double *k0,
                       const int *grid, double cell_size, int inc_self, int es_update) {
    if (!b || num_faces < 0 || (num_faces > 0 && (!face_nodes || !face_type))) return fail(FFH_E_ARG, "null argument");
    if (kind < FFH_SSINT_LENNARD_JONES || kind > FFH_SSINT_STERIC) return fail(FFH_E_ARG, "unknown interaction type");
    if (kind != FFH_SSINT_LENNARD_JONES && num_faces > 0 && !face_opposite) return fail(FFH_E_ARG, "null argument");
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);  // the step gains or loses the interaction
    b->graphs.clear();
    b->free_ssint();
    ffh::ForceInput keep = b->fin;
    std::vector<int32_t> pool_nodes, pool_type, pool_nodes4;
    for (int f = 0; f < num_faces; f++) {
        if (face_type[f] == -1) continue;
        if (face_type[f] < 0 || face_type[f] >= num_types)
            return fail(FFH_E_ARG, "face " + std::to_string(f) + ": interaction type out of range");
        // Face::n[0..3]: the opposite vertex takes steric forces (kinds 2-4) and joins the gather
        for (int k = 0; k < 3; k++) pool_nodes.push_back(face_nodes[3 * f + k]);
        pool_nodes.push_back(kind != FFH_SSINT_LENNARD_JONES ? face_opposite[f] : -1);
        for (int k = 0; k < 3; k++) pool_nodes4.push_back(face_nodes[3 * f + k]);
        pool_nodes4.push_back(face_opposite ? face_opposite[f] : -1);
        pool_type.push_back(face_type[f]);
    }
    const int P = (int)pool_type.size();
    b->fin.ssint_faces.clear();
    if (P == 0) {  // nothing active: no interaction
        int rc = b->upload_forces();
        if (rc) {
            b->fin = keep;
            (void)b->upload_forces();
        }
        return rc;
    }
    if (!grid || (kind != FFH_SSINT_STERIC && (!emin || !rmin || !k0))) return fail(FFH_E_ARG, "null argument");
    if (grid[0] < 1 || grid[1] < 1 || grid[2] < 1 || !(cell_size > 0) || es_update < 1 ||
        (long long)grid[0] * grid[1] * grid[2] > (1 << 28))
        return fail(FFH_E_ARG, "invalid lookup grid or es_update");
    if (num_types < 1) return fail(FFH_E_ARG, "at least one interaction type");
    // steric alone reads no pair parameters (the tables may be null)
    const std::vector<double> none((size_t)num_types * num_types, 0.0);
    if (kind == FFH_SSINT_STERIC) emin = rmin = k0 = none.data();
    for (int t = 0; t < num_types * num_types && kind != FFH_SSINT_STERIC; t++)
        if (!(emin[t] > 0) || !(rmin[t] > 0)) return fail(FFH_E_ARG, "Emin and Rmin must be greater than 0");
    b->fin.ssint_faces = pool_nodes;
    int rc = b->upload_forces();
    if (rc) {
        b->fin = keep;
        (void)b->upload_forces();
        return rc;
    }
    const ffh::Plan &Pl = b->plan;
    std::vector<int32_t> mid_of(Pl.num_nodes_total, -1);  // N_t id -> midpoint record
    for (size_t m = 0; m < Pl.mid_index.size(); m++) mid_of[Pl.mid_index[m]] = (int32_t)m;
    std::vector<int32_t> ends(6 * (size_t)P), blob(P), opp(P, 0);
    for (int p = 0; p < P; p++) {
        if (kind != FFH_SSINT_LENNARD_JONES) {
            const int v = pool_nodes[4 * p + 3];
            if (v < 0 || v >= Pl.num_nodes_total || Pl.lin_of[v] < 0) {
                b->fin = keep;
                (void)b->upload_forces();
                return fail(FFH_E_ARG, "interaction face " + std::to_string(p) + ": the opposite node " +
                                           std::to_string(v) + " is out of range or an edge midpoint");
            }
            opp[p] = Pl.lin_of[v];
        }
        for (int k = 0; k < 3; k++) {  // build_forces has checked every corner is in an element
            const int v = pool_nodes[4 * p + k];
            int a = Pl.lin_of[v], c = a;
            if (a < 0) {
                const int m = mid_of[v];
                a = Pl.lin_of[Pl.mid_a[m]];
                c = Pl.lin_of[Pl.mid_b[m]];
            }
            ends[6 * (size_t)p + 2 * k] = a;
            ends[6 * (size_t)p + 2 * k + 1] = c;
        }
        blob[p] = Pl.slice_blob[ends[6 * (size_t)p] / ffh::kSlice];
    }
    std::vector<int32_t> order(P);
    for (int p = 0; p < P; p++) order[p] = p;
    const int ncell = grid[0] * grid[1] * grid[2];
    int bits = 1;
    while ((1LL << bits) < ncell) bits++;
    SsArgs a = {};
    auto put = [&](const void *h, size_t bytes, void **dst) -> int {
        void *q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(bytes, 8)) != hipSuccess) return fail(FFH_E_NOMEM, "hipMalloc (interactions)");
        b->ss_allocs.push_back(q);
        if (h && hipMemcpy(q, h, bytes, hipMemcpyHostToDevice) != hipSuccess) return fail(FFH_E_HIP, "interaction upload");
        *dst = q;
        return FFH_OK;
    };
    SsCtl c0 = {es_update, es_update, 0, 0};  // World::run starts with a refresh (src/World.cpp:1805)
    rc = put(ends.data(), ends.size() * 4, (void **)&a.ends);
    if (!rc) rc = put(blob.data(), blob.size() * 4, (void **)&a.blob);
    if (!rc) rc = put(pool_type.data(), pool_type.size() * 4, (void **)&a.type);
    if (!rc) rc = put(emin, sizeof(double) * num_types * num_types, (void **)&a.emin);
    if (!rc) rc = put(rmin, sizeof(double) * num_types * num_types, (void **)&a.rmin);
    if (!rc) rc = put(k0, sizeof(double) * num_types * num_types, (void **)&a.k0);
    if (!rc) rc = put(pool_nodes4.data(), pool_nodes4.size() * 4, (void **)&a.nodes4);
    if (!rc) rc = put(opp.data(), opp.size() * 4, (void **)&a.opp);
    if (!rc) rc = put(nullptr, 4 * (size_t)P, (void **)&a.deferred);
    if (!rc) rc = put(nullptr, 8 * sizeof(double) * (size_t)P, (void **)&a.geom);
    if (!rc) rc = put(nullptr, 4 * (size_t)P, (void **)&a.cell);
    if (!rc) rc = put(order.data(), 4 * (size_t)P, (void **)&a.order);
    if (!rc) rc = put(nullptr, 4 * (size_t)P, (void **)&a.sorted);
    if (!rc) rc = put(nullptr, 4 * (size_t)P, (void **)&a.sorted_cell);
    if (!rc) rc = put(nullptr, 4 * (size_t)ncell, (void **)&a.cstart);
    if (!rc) rc = put(nullptr, 4 * (size_t)ncell, (void **)&a.cend);
    if (!rc) rc = put(&c0, sizeof(c0), (void **)&a.ctl);
    if (!rc) rc = put(nullptr, sizeof(double) * (size_t)b->nblobs * b->nblobs, (void **)&b->ss_fe);
    if (!rc) rc = put(nullptr, sizeof(double), (void **)&b->ss_total);
    if (!rc) rc = put(nullptr, 4 * (size_t)P, (void **)&b->ss_cnt);
    if (!rc) rc = put(nullptr, 4 * (size_t)P, (void **)&b->ss_off);
    size_t sb = 0;
    if (!rc && hipcub::DeviceScan::ExclusiveSum(nullptr, sb, b->ss_cnt, b->ss_off, P, b->stream) != hipSuccess)
        rc = fail(FFH_E_HIP, "scan set-up");
    if (!rc) rc = put(nullptr, sb, &b->ss_scan_temp);
    b->ss_scan_bytes = sb;
    size_t tb = 0;
    if (!rc && hipcub::DeviceRadixSort::SortPairs(nullptr, tb, a.cell, a.sorted_cell, a.order, a.sorted, P, 0, bits,
                                                  b->stream) != hipSuccess)
        rc = fail(FFH_E_HIP, "radix sort set-up");
    if (!rc) rc = put(nullptr, tb, &b->ss_temp);
    if (!rc && hipHostMalloc((void **)&b->ss_err_host, sizeof(int), hipHostMallocMapped) != hipSuccess)
        rc = fail(FFH_E_NOMEM, "hipHostMalloc (interaction flag)");
    int *err_dev = nullptr;
    if (!rc && hipHostGetDevicePointer((void **)&err_dev, b->ss_err_host, 0) != hipSuccess)
        rc = fail(FFH_E_HIP, "hipHostGetDevicePointer");
    if (rc) {
        b->free_ssint();
        b->fin = keep;
        (void)b->upload_forces();
        return rc;
    }
    *b->ss_err_host = 0;
    b->ss_err_dev = err_dev;
    a.err = err_dev;
    a.ntypes = num_types;
    a.kind = kind;
    a.st_factor = b->steric_factor;
    a.st_dr = b->steric_dr;
    a.nx = grid[0];
    a.ny = grid[1];
    a.nz = grid[2];
    a.h = cell_size;
    a.inc_self = inc_self ? 1 : 0;
    a.P = P;
    a.ncell = ncell;
    b->ss = a;
    b->ss_temp_bytes = tb;
    b->ss_bits = bits;
    b->ss_on = true;
    return FFH_OK;
}

int ffh_blob_set_steric(ffh_blob *b, double steric_factor, double steric_dr) {
    if (!b) return fail(FFH_E_ARG, "null argument");
    if (!(steric_dr > 0)) return fail(FFH_E_ARG, "steric_dr can only be > 0");
    if (!std::isfinite(steric_factor)) return fail(FFH_E_ARG, "steric_factor must be finite");
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);  // the kernels' arguments change
    b->graphs.clear();
    b->steric_factor = steric_factor;
    b->steric_dr = steric_dr;
    b->ss.st_factor = steric_factor;
    b->ss.st_dr = steric_dr;
    return FFH_OK;
}

}  // extern "C"

// VdW_solver::get_field_energy over the last solve's pairs, with the pair energies and their
// blob-pair keys in pe / pb (room for cap pairs).  grow: the handle's own buffers, enlarged
// with hipMalloc when the pairs do not fit; otherwise more than cap pairs is an error.
static int ssint_energy_impl(ffh_blob *b, double *energy, double *pe, int *pb, size_t cap, bool grow) {
    SsArgs a = b->ss;
    a.pos = b->pos;
    const int P = a.P;
    hipLaunchKernelGGL(k_ss_pair_count, P, 64, 0, b->stream, a, b->ss_cnt);
    HIP_OR_FAIL(hipGetLastError());
    size_t sb = b->ss_scan_bytes;
    HIP_OR_FAIL(hipcub::DeviceScan::ExclusiveSum(b->ss_scan_temp, sb, b->ss_cnt, b->ss_off, P, b->stream));
    int last[2];
    HIP_OR_FAIL(hipMemcpyAsync(&last[0], b->ss_off + P - 1, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipMemcpyAsync(&last[1], b->ss_cnt + P - 1, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    const size_t npairs = (size_t)last[0] + last[1];
    if (npairs > cap) {
        if (!grow)
            return fail(FFH_E_ARG, "the interaction energy has " + std::to_string(npairs) +
                                       " face pairs, more than the no-allocation scratch holds (" +
                                       std::to_string(cap) + ")");
        (void)hipFree(b->ss_pe);
        (void)hipFree(b->ss_pb);
        b->ss_pe = nullptr;
        b->ss_pb = nullptr;
        b->ss_pair_cap = 0;
        const size_t ncap = npairs + npairs / 4 + 1024;
        if (hipMalloc((void **)&b->ss_pe, ncap * sizeof(double)) != hipSuccess ||
            hipMalloc((void **)&b->ss_pb, ncap * sizeof(int)) != hipSuccess)
            return fail(FFH_E_NOMEM, "hipMalloc (interaction pair energies)");
        b->ss_pair_cap = ncap;
        pe = b->ss_pe;
        pb = b->ss_pb;
    }
    if (a.kind == FFH_SSINT_LENNARD_JONES)
        hipLaunchKernelGGL(k_ss_pair_energy<false>, P, 64, 0, b->stream, a, b->nblobs, (const int *)b->ss_off, pe, pb);
    else
        hipLaunchKernelGGL(k_ss_pair_energy<true>, P, 64, 0, b->stream, a, b->nblobs, (const int *)b->ss_off, pe, pb);
    HIP_OR_FAIL(hipGetLastError());
    hipLaunchKernelGGL(k_ss_energy_sum, 1, 64, 0, b->stream, (int)npairs, (const double *)pe, (const int *)pb,
                       b->nblobs, b->ss_fe, b->ss_total);
    HIP_OR_FAIL(hipGetLastError());
    HIP_OR_FAIL(hipMemcpyAsync(energy, b->ss_total, sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    return FFH_OK;
}

// Scratch of ffh_blob_ssint_energy_noalloc: the pair energies (FP64), then their blob-pair
// keys (int32), of at most kNoallocPairs face pairs (12 MiB).  Each GPU holds its own copy
// from the module's load to the end of the process; in host code the name is only a symbol,
// so its device address is looked up once per device (noalloc_scratch) and used instead.
constexpr size_t kNoallocPairs = (size_t)1 << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocPairs * (sizeof(double) + sizeof(int))];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, looked up on the first call for that device with
// that device current (the caller's current device is restored).
int noalloc_scratch(int device, unsigned char **out) {
    if (device < 0 || device >= kMaxDevices) return fail(FFH_E_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        HIP_OR_FAIL(hipGetDevice(&cur));
        if (cur != device) HIP_OR_FAIL(hipSetDevice(device));
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != device) HIP_OR_FAIL(hipSetDevice(cur));
        if (e != hipSuccess) return fail(FFH_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<unsigned char *>(ptr);
    }
    *out = g_scratch_ptr[device];
    return FFH_OK;
}
}  // namespace

extern "C" {

int ffh_blob_ssint_energy(ffh_blob *b, double *energy) {
    if (!b || !energy) return fail(FFH_E_ARG, "null argument");
    if (!b->ss_on) {
        *energy = 0.0;
        return FFH_OK;
    }
    return ssint_energy_impl(b, energy, b->ss_pe, b->ss_pb, b->ss_pair_cap, true);
}

int ffh_blob_ssint_energy_noalloc(ffh_blob *b, double *energy) {
    if (!b || !energy) return fail(FFH_E_ARG, "null argument");
    if (!b->ss_on) {
        *energy = 0.0;
        return FFH_OK;
    }
    unsigned char *ptr = nullptr;
    int rc = noalloc_scratch(b->device, &ptr);
    if (rc) return rc;
    double *pe = reinterpret_cast<double *>(ptr);
    int *pb = reinterpret_cast<int *>(ptr + kNoallocPairs * sizeof(double));
    return ssint_energy_impl(b, energy, pe, pb, kNoallocPairs, false);
}

int ffh_blob_set_box(ffh_blob *b, const double *box, int periodic_mask) {
    if (!b || (periodic_mask && !box)) return fail(FFH_E_ARG, "null argument");
    if (periodic_mask < 0 || periodic_mask > 63) return fail(FFH_E_ARG, "periodic wall mask must be within 0..63");
    if (periodic_mask)
        for (int k = 0; k < 3; k++)
            if (!std::isfinite(box[k])) return fail(FFH_E_ARG, "non-finite box size");
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);  // the step gains or loses the wrap
    b->graphs.clear();
    for (void *q : b->box_allocs) (void)hipFree(q);
    b->box_allocs.clear();
    b->box_mask = 0;
    b->wrap = WrapArgs{};
    if (!periodic_mask) return FFH_OK;
    const ffh::Plan &P = b->plan;
    const int NT = P.num_nodes_total;
    std::vector<int2> ends(NT, make_int2(-1, -1));
    for (int v = 0; v < NT; v++)
        if (P.lin_of[v] >= 0) ends[v] = make_int2(P.lin_of[v], P.lin_of[v]);
    for (size_t k = 0; k < P.mid_index.size(); k++)
        ends[P.mid_index[k]] = make_int2(P.lin_of[P.mid_a[k]], P.lin_of[P.mid_b[k]]);
    for (int v = 0; v < NT; v++)
        if (ends[v].x < 0 || ends[v].y < 0) return fail(FFH_E_ARG, "node " + std::to_string(v) + " is in no element");
    std::vector<int32_t> item_off(b->node_off_h.begin(), b->node_off_h.end()), lo, chunk_off(1, 0);
    for (int k = 0; k < b->nblobs; k++) {
        for (int t = item_off[k]; t < item_off[k + 1]; t += kWrapChunk) lo.push_back(t);
        chunk_off.push_back((int32_t)lo.size());
    }
    lo.push_back(NT);
    const int nc = (int)lo.size() - 1;
    auto put = [&](const void *h, size_t bytes, void **dst) -> int {
        void *q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(bytes, 8)) != hipSuccess) return fail(FFH_E_NOMEM, "hipMalloc (box)");
        b->box_allocs.push_back(q);
        if (h && hipMemcpy(q, h, bytes, hipMemcpyHostToDevice) != hipSuccess) return fail(FFH_E_HIP, "box upload");
        *dst = q;
        return FFH_OK;
    };
    WrapArgs w = {};
    int rc = put(ends.data(), ends.size() * sizeof(int2), (void **)&w.ends);
    if (!rc) rc = put(item_off.data(), item_off.size() * sizeof(int32_t), (void **)&w.blob_item_off);
    if (!rc) rc = put(lo.data(), lo.size() * sizeof(int32_t), (void **)&w.chunk_lo);
    if (!rc) rc = put(chunk_off.data(), chunk_off.size() * sizeof(int32_t), (void **)&w.blob_chunk_off);
    if (!rc) rc = put(nullptr, 6 * (size_t)nc * sizeof(double), (void **)&w.parts);
    if (!rc) rc = put(nullptr, (size_t)b->nblobs * sizeof(double4), (void **)&w.shift);
    if (rc) {
        for (void *q : b->box_allocs) (void)hipFree(q);
        b->box_allocs.clear();
        return rc;
    }
    w.slice_blob = b->slice_blob;
    w.box_x = box[0];
    w.box_y = box[1];
    w.box_z = box[2];
    w.mask = periodic_mask;
    w.N = b->N;
    b->wrap = w;
    b->wrap_chunks = nc;
    b->box_mask = periodic_mask;
    return FFH_OK;
}

int ffh_blob_set_surface_forces(ffh_blob *b, int num_faces, const int *face_nodes, int num_sets, const int *set_off,
                                const int *set_faces, const double *set_force) {
    if (!b || num_faces < 0 || num_sets < 0 || (num_sets > 0 && (!face_nodes || !set_off || !set_faces || !set_force)))
        return fail(FFH_E_ARG, "null argument");
    ffh::ForceInput keep = b->fin;
    if (num_sets == 0) {
        b->fin.faces.clear();
        b->fin.set_off.clear();
        b->fin.set_faces.clear();
        b->fin.set_force.clear();
    } else {
        if (set_off[0] != 0 || set_off[num_sets] < 0) return fail(FFH_E_ARG, "inconsistent surface set offsets");
        for (int i = 0; i < 3 * num_sets; i++)
            if (!std::isfinite(set_force[i])) return fail(FFH_E_ARG, "non-finite surface force");
        b->fin.faces.assign(face_nodes, face_nodes + 3 * (size_t)num_faces);
        b->fin.set_off.assign(set_off, set_off + num_sets + 1);
        b->fin.set_faces.assign(set_faces, set_faces + set_off[num_sets]);
        b->fin.set_force.assign(set_force, set_force + 3 * (size_t)num_sets);
    }
    int rc = b->upload_forces();
    if (rc) {
        b->fin = keep;
        (void)b->upload_forces();
    }
    return rc;
}

int ffh_blob_set_solver(ffh_blob *b, int solver) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    if (solver == 0) return fail(FFH_E_ARG, "the direct solver (FFEA_DIRECT_SOLVER) is not on this path");
    if (solver != FFH_SOLVER_CG && solver != FFH_SOLVER_MASSLUMPED && solver != FFH_SOLVER_CG_NOMASS)
        return fail(FFH_E_ARG, "unknown solver " + std::to_string(solver));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    int rc = FFH_OK;
    if (solver != FFH_SOLVER_CG_NOMASS && !b->rho) rc = b->upload(b->rho, b->plan.matRho);
    if (!rc && solver == FFH_SOLVER_CG && !b->mval) {
        std::vector<double> val, d3, inv;
        std::string msg = ffh::build_mass(b->plan, 1, val, d3, inv);
        if (!msg.empty()) return fail(FFH_E_ARG, msg);
        rc = b->upload(b->mval, val);
        if (!rc) rc = b->upload(b->mdinv, d3);
    }
    if (!rc && solver == FFH_SOLVER_MASSLUMPED && !b->minv) {
        std::vector<double> val, d3, inv;
        std::string msg = ffh::build_mass(b->plan, 2, val, d3, inv);
        if (!msg.empty()) return fail(FFH_E_ARG, msg);
        rc = b->upload(b->minv, inv);
    }
    if (rc) return rc;
    if (solver != b->solver) {
        for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);  // the step's kernels change
        b->graphs.clear();
        b->solver = solver;
        b->plain_count = -1;
        b->cg_chunk = 16;
    }
    return FFH_OK;
}

int ffh_blob_kinetic_energy(ffh_blob *b, double *energy) {
    if (!b || !energy) return fail(FFH_E_ARG, "null argument");
    if (!b->massful()) {
        *energy = 0.0;
        return FFH_OK;
    }
    ElemArgs g = b->elem_args();
    hipLaunchKernelGGL(k_kinetic_energy, b->P_elem, kBlock, 0, b->stream, g, (const double *)b->rho, b->parts_energy());
    HIP_OR_FAIL(hipGetLastError());
    std::vector<double> h(b->P_elem);
    HIP_OR_FAIL(hipMemcpyAsync(h.data(), b->parts_energy(), h.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    double s = 0.0;
    for (double v : h) s += v;
    *energy = 0.5 * s;
    return FFH_OK;
}

int ffh_blob_get_rhs(ffh_blob *b, double *f) {
    if (!b || !f) return fail(FFH_E_ARG, "null argument");
    return get_node_vector(b, b->f, f);
}

int ffh_blob_get_rng_state(ffh_blob *b, unsigned long long *state6) {
    if (!b || !state6) return fail(FFH_E_ARG, "null argument");
    uint32_t s[6];
    HIP_OR_FAIL(hipMemcpyAsync(s, b->rng_state, sizeof(s), hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    for (int i = 0; i < 6; i++) state6[i] = s[i];
    return FFH_OK;
}

int ffh_blob_set_rng_state(ffh_blob *b, const unsigned long long *state6) {
    if (!b || !state6) return fail(FFH_E_ARG, "null argument");
    uint32_t s[6];
    for (int i = 0; i < 6; i++) {
        if (state6[i] >= (i < 3 ? ffh::kM1 : ffh::kM2)) return fail(FFH_E_ARG, "RngStream state out of range");
        s[i] = (uint32_t)state6[i];
    }
    HIP_OR_FAIL(hipMemcpyAsync(b->rng_state, s, sizeof(s), hipMemcpyHostToDevice, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    return FFH_OK;
}

int ffh_blob_last_cg(ffh_blob *b, int *iterations, double *residual2) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    if (iterations) *iterations = b->last_iters;
    if (residual2) *residual2 = b->last_res;
    return FFH_OK;
}

int ffh_blob_sizes(ffh_blob *b, int *num_linear, int *num_elements, long long *num_blocks, long long *num_slots) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    if (num_linear) *num_linear = b->plan.num_real_linear;
    if (num_elements) *num_elements = b->E;
    if (num_blocks) *num_blocks = b->plan.num_blocks;
    if (num_slots) *num_slots = b->plan.num_slots;
    return FFH_OK;
}

void *ffh_blob_stream(ffh_blob *b) { return b ? (void *)b->stream : nullptr; }

int ffh_blob_synchronize(ffh_blob *b) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    return FFH_OK;
}

int ffh_blob_set_profiling(ffh_blob *b, int enable) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    b->collect_events(true);
    b->profiling = enable != 0;
    for (int k = 0; k < FFH_NUM_KERNELS; k++) {
        b->ktime[k] = 0;
        b->kcount[k] = 0;
    }
    return FFH_OK;
}

int ffh_blob_kernel_times(ffh_blob *b, double *ms, long long *counts) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    b->collect_events(true);
    for (int k = 0; k < FFH_NUM_KERNELS; k++) {
        if (ms) ms[k] = b->ktime[k];
        if (counts) counts[k] = b->kcount[k];
    }
    return FFH_OK;
}

// ---- learned preconditioner -----------------------------------------------------------------

static int mp_ready_check(ffh_blob *b) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    if (!b->mp_ready) return fail(FFH_E_ARG, "no network weights: call ffh_mpnn_init or ffh_mpnn_set_weights");
    return FFH_OK;
}

int ffh_mpnn_num_params(void) { return kNumParams; }

int ffh_mpnn_set_weights(ffh_blob *b, const float *w) {
    if (!b || !w) return fail(FFH_E_ARG, "null argument");
    for (int i = 0; i < kNumParams; i++)
        if (!std::isfinite(w[i])) return fail(FFH_E_ARG, "non-finite network weight");
    int rc;
    if ((rc = b->mp_alloc())) return rc;
    HIP_OR_FAIL(hipMemcpyAsync(b->mp_w, w, kNumParams * sizeof(float), hipMemcpyHostToDevice, b->stream));
    HIP_OR_FAIL(hipMemsetAsync(b->mp_m1, 0, kNumParams * sizeof(float), b->stream));
    HIP_OR_FAIL(hipMemsetAsync(b->mp_m2, 0, kNumParams * sizeof(float), b->stream));
    b->adam_t = 0;
    if ((rc = b->mp_pack())) return rc;
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    b->mp_ready = true;
    return FFH_OK;
}

int ffh_mpnn_get_weights(ffh_blob *b, float *w) {
    if (int rc = mp_ready_check(b)) return rc;
    if (!w) return fail(FFH_E_ARG, "null argument");
    HIP_OR_FAIL(hipMemcpyAsync(w, b->mp_w, kNumParams * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    return FFH_OK;
}

int ffh_mpnn_init(ffh_blob *b, unsigned long long seed) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    // He-normal hidden weights, zero biases, zero output layer (the untrained network is Jacobi)
    std::mt19937_64 gen(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<float> w(kNumParams, 0.0f);
    for (int i = 0; i < kHid * kNIn; i++) w[kOffWe + i] = (float)(nd(gen) * std::sqrt(2.0 / kNIn));
    for (int l = 0; l < kNLay; l++)
        for (int i = 0; i < 2 * kHid * kHid; i++) w[kOffLay + l * kLayStride + i] = (float)(nd(gen) * std::sqrt(1.0 / kHid));
    return ffh_mpnn_set_weights(b, w.data());
}

int ffh_blob_set_preconditioner(ffh_blob *b, int kind, int fallback_iterations) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    if (kind != FFH_PRECOND_JACOBI && kind != FFH_PRECOND_LEARNED) return fail(FFH_E_ARG, "unknown preconditioner");
    if (kind == FFH_PRECOND_LEARNED && !b->mp_ready) return mp_ready_check(b);
    if (fallback_iterations < 0) return fail(FFH_E_ARG, "negative fallback count");
    b->precond = kind;
    b->fallback_iters = fallback_iterations;
    return FFH_OK;
}

int ffh_mpnn_apply(ffh_blob *b, const double *rin, double *zout) {
    if (int rc = mp_ready_check(b)) return rc;
    if (!rin || !zout) return fail(FFH_E_ARG, "null argument");
    const ffh::Plan &P = b->plan;
    std::vector<double> rl(3 * (size_t)b->N, 0.0), zl(3 * (size_t)b->N);
    for (int i = 0; i < b->N; i++)
        if (P.orig_of[i] >= 0 && !P.pinned[i])
            for (int c = 0; c < 3; c++) rl[3 * i + c] = rin[3 * (size_t)P.orig_of[i] + c];
    HIP_OR_FAIL(hipMemcpyAsync(b->tmp_in, rl.data(), rl.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
    // u = D^-1 r into z, partials of |u|^2 -- the update kernel's work, here on a given r
    hipLaunchKernelGGL(k_mpnn_u, P.num_slices, 192, 0, b->stream, P.num_slices, (const double *)b->dinv,
                       (const double *)b->tmp_in, b->z, b->parts_uu());
    HIP_OR_FAIL(hipGetLastError());
    if (int rc = b->totals(1 << kTotUU)) return rc;
    if (int rc = b->mpnn_apply<double>(b->tmp_in, nullptr, false)) return rc;
    HIP_OR_FAIL(hipMemcpyAsync(zl.data(), b->z, zl.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    std::memset(zout, 0, sizeof(double) * 3 * (size_t)P.num_nodes_total);
    for (int i = 0; i < b->N; i++)
        if (P.orig_of[i] >= 0)
            for (int c = 0; c < 3; c++) zout[3 * (size_t)P.orig_of[i] + c] = zl[3 * i + c];
    return FFH_OK;
}

int ffh_mpnn_harvest(ffh_blob *b, int max_snapshots) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    if (max_snapshots < 0) return fail(FFH_E_ARG, "negative snapshot count");
    int rc;
    if ((rc = b->mp_alloc())) return rc;
    if (max_snapshots > b->hv_cap) {
        const size_t n3 = 3 * (size_t)b->N;
        float *r = nullptr, *x = nullptr;
        if ((rc = b->alloc(r, n3 * max_snapshots)) || (rc = b->alloc(x, n3 * max_snapshots))) return rc;
        b->hv_r = r;  // the old (smaller) buffers stay allocated until the handle is destroyed
        b->hv_x = x;
        b->hv_cap = max_snapshots;
    }
    b->hv_on = max_snapshots > 0;
    b->hv_count = 0;
    b->hv_pairs = 0;
    return FFH_OK;
}

int ffh_mpnn_harvest_count(ffh_blob *b, int *snapshots, long long *pairs) {
    if (!b) return fail(FFH_E_ARG, "null blob");
    if (snapshots) *snapshots = b->hv_count;
    if (pairs) *pairs = b->hv_pairs;
    return FFH_OK;
}

int ffh_mpnn_train(ffh_blob *b, int epochs, float lr, unsigned long long seed, float *loss_per_epoch) {
    if (int rc = mp_ready_check(b)) return rc;
    if (epochs < 1 || epochs > 4096 || !(lr > 0)) return fail(FFH_E_ARG, "invalid training parameters");
    if (b->hv_count < 1) return fail(FFH_E_ARG, "no harvested snapshots: call ffh_mpnn_harvest, then solve");
    b->hv_on = false;  // the training data are fixed from here on
    HIP_OR_FAIL(hipMemsetAsync(b->mp_loss, 0, epochs * sizeof(float), b->stream));
    std::mt19937_64 gen(seed);
    std::vector<int> order(b->hv_count);
    for (int i = 0; i < b->hv_count; i++) order[i] = i;
    const size_t n3 = 3 * (size_t)b->N;
    for (int e = 0; e < epochs; e++) {
        std::shuffle(order.begin(), order.end(), gen);
        for (int s : order) {
            int rc;
            if ((rc = b->mpnn_forward_backward(b->hv_r + (size_t)s * n3, b->hv_x + (size_t)s * n3, b->mp_loss + e)))
                return rc;
            if ((rc = b->mpnn_adam(lr))) return rc;
        }
    }
    std::vector<float> h(epochs);
    HIP_OR_FAIL(hipMemcpyAsync(h.data(), b->mp_loss, epochs * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    if (loss_per_epoch)
        for (int e = 0; e < epochs; e++) loss_per_epoch[e] = h[e] / (float)b->hv_count;
    return FFH_OK;
}

int ffh_mpnn_loss_grad(ffh_blob *b, const double *rin, const double *ein, double *loss, float *grad) {
    if (int rc = mp_ready_check(b)) return rc;
    if (!rin || !ein) return fail(FFH_E_ARG, "null argument");
    const ffh::Plan &P = b->plan;
    const size_t n3 = 3 * (size_t)b->N;
    std::vector<float> rl(n3, 0.0f), el(n3, 0.0f);
    for (int i = 0; i < b->N; i++)
        if (P.orig_of[i] >= 0 && !P.pinned[i])
            for (int c = 0; c < 3; c++) {
                rl[3 * i + c] = (float)rin[3 * (size_t)P.orig_of[i] + c];
                el[3 * i + c] = (float)ein[3 * (size_t)P.orig_of[i] + c];
            }
    float *dr = reinterpret_cast<float *>(b->tmp_in), *de = reinterpret_cast<float *>(b->tmp_out);
    HIP_OR_FAIL(hipMemcpyAsync(dr, rl.data(), n3 * sizeof(float), hipMemcpyHostToDevice, b->stream));
    HIP_OR_FAIL(hipMemcpyAsync(de, el.data(), n3 * sizeof(float), hipMemcpyHostToDevice, b->stream));
    HIP_OR_FAIL(hipMemsetAsync(b->mp_loss + 4095, 0, sizeof(float), b->stream));
    if (int rc = b->mpnn_forward_backward(dr, de, b->mp_loss + 4095)) return rc;
    float lh = 0.0f;
    HIP_OR_FAIL(hipMemcpyAsync(&lh, b->mp_loss + 4095, sizeof(float), hipMemcpyDeviceToHost, b->stream));
    if (grad) HIP_OR_FAIL(hipMemcpyAsync(grad, b->mp_g, kNumParams * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    if (loss) *loss = lh;
    return FFH_OK;
}

int ffh_blob_strain_energy(ffh_blob *b, double *energy) {
    if (!b || !energy) return fail(FFH_E_ARG, "null argument");
    ElemArgs g = b->elem_args();
    hipLaunchKernelGGL(k_strain_energy, b->P_elem, kBlock, 0, b->stream, g, b->parts_energy());
    HIP_OR_FAIL(hipGetLastError());
    std::vector<double> h(b->P_elem);
    HIP_OR_FAIL(hipMemcpyAsync(h.data(), b->parts_energy(), h.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_OR_FAIL(hipStreamSynchronize(b->stream));
    double s = 0.0;
    for (double v : h) s += v;
    *energy = 0.5 * s;
    return FFH_OK;
}

}  // extern "C"
