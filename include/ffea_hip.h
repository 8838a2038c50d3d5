/*
 * ffea_hip.h -- C-ABI of the MI355X (gfx950) implementation of FFEA's massless
 * (CG_nomass) fluctuating-FEM time step.
 *
 * One handle = one dynamic blob plus the RngStream that feeds it, resident on one
 * GPU with its own HIP stream.  A handle is single-owner and not reentrant.
 * Every call returns FFH_OK (0) or a negative FFH_E* code; ffh_last_error() gives
 * the message of the last failure on the calling thread.
 *
 * Host arrays use FFEA's full node numbering (N_t nodes: linear nodes and the
 * quadratic edge midpoints of the 10-node topology), positions and radii in meso
 * units after `scale`, materials in SI as written in a .mat file.  On the device
 * the blob is compacted to its linear nodes (midpoint rows of the reference's
 * matrix are identity and carry no force; their positions are edge averages).
 */
#ifndef FFEA_HIP_H
#define FFEA_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define FFH_OK 0
#define FFH_E_ARG -1         /* invalid argument or unsupported mesh */
#define FFH_E_HIP -2         /* HIP runtime failure */
#define FFH_E_INVERSION -3   /* an element changed orientation (run must stop) */
#define FFH_E_CG -4          /* CG did not converge in max_iterations_cg */
#define FFH_E_NOMEM -5

#define FFH_PRECISION_FP64 0  /* element, assembly, SpMV and CG all FP64 */
#define FFH_PRECISION_MIXED 1 /* FP32 element/assembly/SpMV, FP64 CG scalars, RHS and integration */

typedef struct ffh_blob ffh_blob;

/* Message of the last failed call on this thread ("" if none). */
const char *ffh_last_error(void);

/* Library version, (major << 16) | minor. */
int ffh_version(void);

/*
 * Create a blob on the current HIP device.
 * Replaces Blob::init's post-load setup (src/Blob.cpp:335-412: linearise_element,
 * calc_rest_state_info :3760-3822, calculate_node_element_connectivity :3974-4019,
 * stokes_drag :408-412) and Solver::init (include/Solver.h:45;
 * NoMassCGSolver::init, src/NoMassCGSolver.cpp:63-162).
 *   num_nodes        N_t, all nodes of the .node file
 *   num_elements     E
 *   pos              [3*N_t] rest positions, meso units (Blob::load_nodes, :2480-2495)
 *   topo             [10*E] node indices per element, FFEA order (load_topology :2606-2622)
 *   material_si      [6*E] density, shear_visc, bulk_visc, shear_mod, bulk_mod, dielectric
 *                    (load_material_params, src/Blob.cpp:2821-2876)
 *   stokes_radius    [N_t] meso units (load_stokes_params :2880-2947); must be 0 on midpoints
 *   pinned           [num_pinned] node indices (load_pinned_nodes :3691-3755)
 *   dt, kT, epsilon2, stokes_visc   meso units as SimulationParams holds them
 *   rng_state        [6] RngStream Cg state the blob's stream starts from (World::init
 *                    RNG.2/.3, src/World.cpp:287-323; or a .fcp checkpoint line)
 *   precision        FFH_PRECISION_FP64 or FFH_PRECISION_MIXED
 * Returns NULL on failure (see ffh_last_error).
 */
ffh_blob *ffh_blob_create(int num_nodes, int num_elements, const double *pos, const int *topo,
                          const double *material_si, const double *stokes_radius, int num_pinned,
                          const int *pinned, double dt, double kT, double epsilon2,
                          int max_iterations_cg, int calc_noise, int calc_stokes, double stokes_visc,
                          const unsigned long long *rng_state, int precision);

/*
 * ffh_blob_create with the materials as tetra_element_linear holds them after
 * Blob::load_material_params (src/Blob.cpp:2861-2866): material_meso [5*E] = rho, A (shear
 * viscosity), B (second viscosity coefficient), G (shear modulus), E (bulk modulus), meso units.
 * For callers that hold converted elements, such as a Solver::init (include/Solver.h:45), so
 * that no conversion back to SI and forth rounds them again.  Everything else as above.
 */
ffh_blob *ffh_blob_create_meso(int num_nodes, int num_elements, const double *pos, const int *topo,
                               const double *material_meso, const double *stokes_radius, int num_pinned,
                               const int *pinned, double dt, double kT, double epsilon2,
                               int max_iterations_cg, int calc_noise, int calc_stokes, double stokes_visc,
                               const unsigned long long *rng_state, int precision);

/*
 * Create a batch: several blobs of one World on the current HIP device, stepped together (one
 * kernel launch per phase for all of them) and fed by one RngStream, in World::run's order
 * (src/World.cpp:2020-2041): the element phases of blobs 0..B-1 (7 draws per element), then
 * their position phases (3 draws per node).  With one OpenMP thread this is the reference's
 * draw order for params.num_blobs > 1 (World::init creates the streams,
 * src/World.cpp:298-317, and hands the same array to every Blob::config, :2618-2622).
 *   num_blobs                 B
 *   num_nodes[B], num_elements[B], num_pinned[B]   per-blob counts
 *   pos, stokes_radius        blobs' node arrays concatenated (blob 0's N_t nodes first)
 *   topo, material_si         blobs' element arrays concatenated; topology and pinned indices
 *   pinned                    are local to their blob, as in each blob's .top / .pin file
 * Every other argument is as for ffh_blob_create (SimulationParams is shared by the World).
 * The handle works with every ffh_blob_* call; node arrays are then the concatenation over
 * blobs, and each blob's CG runs on its own scalars and stops on its own test.
 */
ffh_blob *ffh_batch_create(int num_blobs, const int *num_nodes, const int *num_elements, const double *pos,
                           const int *topo, const double *material_si, const double *stokes_radius,
                           const int *num_pinned, const int *pinned, double dt, double kT, double epsilon2,
                           int max_iterations_cg, int calc_noise, int calc_stokes, double stokes_visc,
                           const unsigned long long *rng_state, int precision);

/* Number of blobs of a handle (1 for ffh_blob_create). */
int ffh_batch_num_blobs(ffh_blob *b);

/* Per-blob result of the last CG solve: iterations, final ||r||^2/||f||^2 and FFH_OK or
 * FFH_E_CG (then that blob's positions were not integrated), arrays of num_blobs entries
 * (NoMassCGSolver::solve's outcome per Blob, src/NoMassCGSolver.cpp:165-197). */
int ffh_batch_last_cg(ffh_blob *b, int *iterations, double *residual2, int *status);

/* Release all device memory of the blob (Blob::~Blob / NoMassCGSolver::~NoMassCGSolver). */
void ffh_blob_destroy(ffh_blob *b);

/*
 * Blob::update_internal_forces (src/Blob.cpp:638-735; include/Blob.h:130): the element
 * loop (Jacobian, shape gradients, volume + inversion check, Kelvin-Voigt viscosity
 * matrix, elastic + fluctuating stress, nodal forces) followed by the assembly of the
 * viscosity matrix into its fixed pattern (SparseMatrixFixedPattern::build,
 * src/SparseMatrixFixedPattern.cpp:85-93).  Returns FFH_OK, or FFH_E_INVERSION with the
 * number of inverted elements in *num_inversions (may be NULL).
 */
int ffh_blob_update_internal_forces(ffh_blob *b, int *num_inversions);

/*
 * Blob::update_positions (src/Blob.cpp:737-764): aggregate_forces_and_solve (:3826-3932:
 * force gather, Stokes noise, pinned rows, NoMassCGSolver::solve) then euler_integrate
 * (:3937-3953) and linearise_element.  Returns FFH_OK or FFH_E_CG (then, as in the
 * reference, positions are not integrated).
 */
int ffh_blob_update_positions(ffh_blob *b);

/*
 * nsteps iterations of World::run for this blob (src/World.cpp:1811-2043): element phase,
 * then position phase, per step.  Stops at the first error.  Asynchronous with respect to
 * the host except for the convergence checks of the CG solve.
 */
int ffh_blob_step(ffh_blob *b, int nsteps);

/* Launch shapes, tuned at creation by timing this blob's own kernels (set FFH_AUTOTUNE=0 in
 * the environment to keep the defaults 256 / 512 / 2 / 3 / 0): threads per workgroup of the
 * element kernel (64, 128 or 256); threads per workgroup of the assembly (256, 512 or 1024: one
 * workgroup per 64-row slice, whose first wave sums the diagonal blocks) and the 64-slot tiles of
 * upper blocks each other assembly wave sums per pass (1, 2 or 4; 1 or 2 with 1024 threads);
 * waves per 64-row SpMV slice (3: one per component; 1: one wave computes all three) and
 * non-temporal matrix loads in the SpMV (1: the streamed matrix leaves the caches to the
 * gathered vector; wins when it streams from HBM, loses when it fits the Infinity Cache).
 * Results are bit-identical for every choice. */
int ffh_blob_tuning(ffh_blob *b, int *element_block, int *assembly_block, int *assembly_tiles,
                    int *spmv_waves_per_slice, int *spmv_nontemporal);
int ffh_blob_set_tuning(ffh_blob *b, int element_block, int assembly_block, int assembly_tiles,
                        int spmv_waves_per_slice, int spmv_nontemporal);

/*
 * Forces from outside the element loop -- surface (face) forces of interactions, springs,
 * constant forces -- as computed by the caller: [3*N_t] host array in meso units, or NULL to
 * clear.  They are added to the force vector of every following position phase after the
 * element forces and before the Stokes noise, as Blob::aggregate_forces_and_solve adds face
 * forces (src/Blob.cpp:3842-3851); forces on edge midpoints are moved to the edge ends as
 * linearise_force does (src/Blob.cpp:1066-1085), and pinned rows stay zero.  The caller
 * refreshes them as often as its interaction model needs.  Additions are grouped as one
 * pre-summed term per node, so the force vector can differ from the reference's in the last bit.
 */
int ffh_blob_set_external_forces(ffh_blob *b, const double *force);

/* ffh_blob_step replays one captured hipGraph per step (element, assembly, force gather and
 * the CG iterations of the last solve's count; the host polls once per step) unless disabled
 * here, or the learned preconditioner, harvesting or profiling is on.  Results are the same
 * either way: the same kernels run in the same order.  Enabled by default. */
int ffh_blob_set_graphs(ffh_blob *b, int enable);

/* Number of position phases completed by this handle (update_positions and ffh_blob_step);
 * after ffh_blob_step returns FFH_E_CG it tells how many steps were made, the failing one
 * included (World::run goes on after a failed solve, src/World.cpp:2042). */
int ffh_blob_steps_done(ffh_blob *b, long long *steps);

/*
 * Solver::solve (include/Solver.h:51; NoMassCGSolver::solve, src/NoMassCGSolver.cpp:165-197):
 * x is [3*N_t] host memory, the force on entry, the velocity on exit.  Uses the viscosity
 * matrix of the last update_internal_forces (or of creation).  With a mass-ful solver
 * (ffh_blob_set_solver) the mass matrix, and x is the acceleration on exit.
 */
int ffh_blob_solve(ffh_blob *b, double *x);

/*
 * Solver::apply_matrix (include/Solver.h:54): result = K * in over [3*N_t] host vectors
 * (K = assembled viscosity matrix with Stokes drag; midpoint and pinned rows identity).
 * Massless solver only (FFH_E_ARG otherwise).
 */
int ffh_blob_apply_matrix(ffh_blob *b, const double *in, double *result);

/*
 * The build step of NoMassCGSolver::solve (SparseMatrixFixedPattern::build,
 * src/SparseMatrixFixedPattern.cpp:85-93, with calc_inverse_diagonal): the viscosity matrix and
 * its inverse diagonal assembled at the handle's current positions, without the element
 * loop's forces or draws.  ffh_blob_solve and ffh_blob_apply_matrix then use it.  Massless
 * solver only.  ffh_blob_solve / ffh_blob_apply_matrix stage their vectors through pinned host
 * buffers of 3 N doubles (N linear nodes), allocated on first use.
 */
int ffh_blob_update_matrix(ffh_blob *b);

/*
 * Element stress magnitudes for Blob::make_stress_measurements (src/Blob.cpp:1540-1550): with
 * the output enabled, every element phase also stores internal_stress_mag per element
 * (src/Blob.cpp:700: sqrt(mat3_double_contraction_symmetric(stress)) of the elastic + bulk +
 * fluctuating stress, src/mat_vec_fns.cpp:131-134), E values in the handle's element order;
 * ffh_blob_get_stress_magnitudes copies the last ones out (meso pressure units).
 */
int ffh_blob_set_stress_output(ffh_blob *b, int enable);
int ffh_blob_get_stress_magnitudes(ffh_blob *b, double *out);

/*
 * Forces from outside the element loop, computed on the device every step from the current
 * positions and added where the reference adds them.  Node ids are in the handle's N_t
 * numbering (a batch: its blobs' nodes concatenated), forces in meso units.  Each call
 * replaces that kind of force; a count of 0 clears it.
 *
 * ffh_blob_set_springs: World::apply_springs (src/World.cpp:3808-3846) -- spring i pulls its
 *   ends node0[i], node1[i] (vertices, possibly in different blobs of a batch) with k (|d| - l)
 *   along their separation d; k and l as World::load_springs converts them (:3509-3510).  Added
 *   to the zeroed force vector before the constant and element forces.
 * ffh_blob_spring_energy: World::make_measurements' spring energy (:4316-4329): sum of
 *   k (|d| - l)^2 / 2, summed per blob pair in spring order, then over the pairs.
 * ffh_blob_set_constant_forces: Blob::apply_ctforces' linear forces (src/Blob.cpp:2096-2103),
 *   force[3i..3i+2] on node[i] in the order Blob::load_ctforces stores them (:3184-3241);
 *   added after the springs, before the element forces.  On an edge midpoint the force reaches
 *   the edge ends as linearise_force passes it on (:1066-1085).
 * ffh_blob_set_rotational_forces: Blob::apply_ctforces' rotational forces (CTF 2, :2104-2136)
 *   with a point + direction axis ('p' types, as Blob::load_ctforces stores them, :3246-3346):
 *   axis[6i..6i+5] the axis point (meso length) and unit direction, F[i] in meso force,
 *   torque[i] 1 for 't', 0 for 'f'.  Every step, with r from the node to its foot on the axis,
 *   the node gets F (r x d) ('f') or F/|r| (r x d) ('t'), after the linear constant forces;
 *   on an edge midpoint evaluated at the midpoint and passed on like a linear force.
 * ffh_blob_set_surface_forces: Blob::apply_ctforces' linear surface forces (:2137-2159) --
 *   face_nodes [3*num_faces] are the surface faces' vertices (Face::n[0..2]); set s is the
 *   faces set_faces[set_off[s] .. set_off[s+1]) with total force set_force[3s..3s+2], shared
 *   over its faces in proportion to their current areas, a third to each vertex; face forces
 *   are added face by face after the element forces (:3842-3851).
 * The interaction solvers (VdW / steric / LJ, PreComp) are not on this path.
 */
int ffh_blob_set_springs(ffh_blob *b, int num_springs, const int *node0, const int *node1, const double *k,
                         const double *l);
int ffh_blob_spring_energy(ffh_blob *b, double *energy);
int ffh_blob_set_constant_forces(ffh_blob *b, int num_forces, const int *node, const double *force);
int ffh_blob_set_rotational_forces(ffh_blob *b, int num_forces, const int *node, const double *F, const double *axis,
                                   const int *torque);
int ffh_blob_set_surface_forces(ffh_blob *b, int num_faces, const int *face_nodes, int num_sets, const int *set_off,
                                const int *set_faces, const double *set_force);

/*
 * Surface-surface interactions (VdW_solver::solve, do_lj_interaction, do_gensoft_interaction,
 * src/VdW_solver.cpp:158-384, 536-831; SSINT_matrix, src/LJ_matrix.cpp:122-199; the lookup grid
 * of NearestNeighbourLinkedListCube, src/NearestNeighbourLinkedListCube.cpp:26-77).
 * kind: FFH_SSINT_LENNARD_JONES, FFH_SSINT_LJSTERIC (Lennard-Jones with the soft core below
 * Rmin, calc_ljinterpolated_force_pair_matrix), FFH_SSINT_GENSOFT (rmin = the cutoff, k0) or
 * FFH_SSINT_STERIC (the volume exclusion alone, Steric_solver; emin, rmin, k0 unread, may be
 * null).
 * face_nodes [3*num_faces]: the surface faces' corners (N_t ids of the handle, blobs in order,
 * each blob's faces in .surf order); face_opposite [num_faces]: Face::n[3], the element vertex
 * opposite each face (for every kind but lennard-jones, for which it may be null);
 * face_type [num_faces]: the .vdw / .ssint type, -1 for an inactive face;
 * emin, rmin, k0 [num_types^2] in meso units (Emin * area^2 / Energy, Rmin / length,
 * k0 * area / Energy), entry type1 * num_types + type2; grid [3] = es_N_x/y/z and cell_size =
 * ssint_cutoff (meso) of the lookup grid; inc_self_ssint; es_update (face geometry and grid
 * refreshed on the first step and every es_update steps, as World::run does).  Every step
 * then starts with the solve: each pair of active faces in neighbouring cells (i < j in pool
 * order, other blob unless inc_self, normals not facing the same way; for the kinds other than
 * lennard-jones also faces in front of each other and no shared node within a blob) interacts.
 * For ljsteric, gensoft and steric, where the two face elements (corners and opposite vertex)
 * overlap (Face::checkTetraIntersection, src/CheckTetrahedraOverlap.cpp) with a nonzero common
 * volume V (src/VolumeIntersection.cpp), the pair gets the steric forces instead
 * (VdW_solver::do_steric_interaction, src/VdW_solver.cpp:458-487; Face.cpp:473-514): energy
 * steric_factor * V and forces steric_factor * phi_j dV/dr on the four nodes of each face
 * (ffh_blob_set_steric); otherwise the three-point-quadrature forces of its potential (none for
 * steric) on the corners.  The faces' forces reach the nodes with the other face forces after
 * the element forces (src/Blob.cpp:3842-3851).  All faces inactive (or num_faces 0) removes
 * the interaction.  Linear surface forces (CTF 3) together with it are refused.  A face
 * centroid outside the grid after one period fails the step (FFH_E_ARG), as
 * LinkedListCube::add_node_to_stack does; so does a face element whose geometry makes the
 * reference throw (arr3Normalise2 of a zero vector).
 * ffh_blob_ssint_energy: VdW_solver::get_field_energy(-1, -1) of the last solve (meso energy),
 * summed in the solve's order; call it between ffh_blob_update_internal_forces and
 * ffh_blob_update_positions, as World::make_measurements does.
 */
#define FFH_SSINT_LENNARD_JONES 1
#define FFH_SSINT_LJSTERIC 2
#define FFH_SSINT_GENSOFT 3
#define FFH_SSINT_STERIC 4
int ffh_blob_set_ssint(ffh_blob *b, int kind, int num_faces, const int *face_nodes, const int *face_opposite,
                       const int *face_type, int num_types, const double *emin, const double *rmin, const double *k0,
                       const int *grid, double cell_size, int inc_self, int es_update);

/*
 * The steric parameters SimulationParams reads as steric_factor and steric_dr (aliases
 * vdw_steric_factor, vdw_steric_dr; src/SimulationParams.cpp:412-418, 749-755), used as given
 * (meso units, no conversion).  Defaults 1 and 5e-3 (src/SimulationParams.cpp:44-45).
 * steric_dr must be > 0 (FFH_E_ARG otherwise, as the reference refuses it).
 */
int ffh_blob_set_steric(ffh_blob *b, double steric_factor, double steric_dr);
int ffh_blob_ssint_energy(ffh_blob *b, double *energy);

/*
 * ffh_blob_ssint_energy without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  ffh_blob_ssint_energy enlarges the handle's pair-energy
 * buffers with hipMalloc when a solve has more face pairs than before; this variant instead
 * keeps the pair energies and keys in one library-owned device array per GPU (12 MiB, held
 * until the process ends), calls no hipMalloc / hipFree, and returns FFH_E_ARG when the last
 * solve has more than 1,048,576 face pairs.  The result equals ffh_blob_ssint_energy's bit for
 * bit.  Every call on a GPU shares that GPU's array: calls that overlap on one GPU (two
 * threads, or two handles on different streams) race and must be serialised by the caller.
 */
int ffh_blob_ssint_energy_noalloc(ffh_blob *b, double *energy);

/*
 * World::run's periodic box (src/World.cpp:1850-1901; box_dim from World::init, :455-471).
 * With periodic_mask != 0, every step (ffh_blob_step) and every ffh_blob_update_internal_forces
 * begins as the reference's step does: for each blob, the centroid of its N_t nodes
 * (Blob::get_centroid, src/Blob.cpp:971-983); per axis, below 0 with that low wall periodic,
 * else above box[axis] with that high wall periodic, the blob moves by +/- box[axis]
 * (Blob::move, :935-949).  periodic_mask bits 0..5: wall_x_1, wall_x_2, wall_y_1, wall_y_2,
 * wall_z_1, wall_z_2 (SimulationParams: PBC is the default for all six); 0 removes the box.
 * Hard walls (Blob::enforce_box_boundaries) are not on this path.
 */
int ffh_blob_set_box(ffh_blob *b, const double *box, int periodic_mask);

/*
 * Mass-ful solvers: Blob::init's solver choice (src/Blob.cpp:358-373) with the codes of
 * include/FFEA_return_codes.h:30-33.  FFH_SOLVER_CG is ConjugateGradientSolver on the
 * consistent mass matrix (src/ConjugateGradientSolver.cpp:62-230), FFH_SOLVER_MASSLUMPED is
 * MassLumpedSolver (src/MassLumpedSolver.cpp:40-90), FFH_SOLVER_CG_NOMASS (the default) the
 * massless NoMassCGSolver.  With a mass-ful solver the element phase adds the viscous force
 * K_e u_e of the node velocities (src/Blob.cpp:703-706), the force phase the Stokes drag
 * -zeta v (:3854-3877), the solve gives accelerations and the integration is vel += a dt,
 * pos += vel dt (:3955-3967); the viscosity matrix is not assembled.  The direct solver
 * (FFEA_DIRECT_SOLVER) is not on this path: FFH_E_ARG.  Velocities start at zero.
 */
#define FFH_SOLVER_CG 1
#define FFH_SOLVER_MASSLUMPED 2
#define FFH_SOLVER_CG_NOMASS 3
int ffh_blob_set_solver(ffh_blob *b, int solver);
/* Velocities of all N_t nodes [3*N_t], meso units (a mass-ful restart).  Midpoint velocities
 * are kept as given and reported back: the solve gives midpoints zero acceleration, and their
 * drag is zero on this path, so they never change and never act. */
int ffh_blob_set_velocities(ffh_blob *b, const double *vel);
/* Kinetic energy of Blob::make_measurements (src/Blob.cpp:1385-1407, :1431), summed over the
 * handle's blobs, meso energy units; 0 with the massless solver. */
int ffh_blob_kinetic_energy(ffh_blob *b, double *energy);

/* Node state over all N_t nodes, host memory [3*N_t], meso units.  get_positions re-averages
 * midpoints (linearise_element); get_velocities reports 0 on midpoints, like the reference. */
int ffh_blob_set_positions(ffh_blob *b, const double *pos);
int ffh_blob_get_positions(ffh_blob *b, double *pos);
int ffh_blob_get_velocities(ffh_blob *b, double *vel);

/* The force vector handed to the last solve (after gather, noise and pinning), [3*N_t]. */
int ffh_blob_get_rhs(ffh_blob *b, double *f);

/* RngStream Cg state (RngStream::GetState / SetSeed, src/RngStream.cpp:388-395, :441-445),
 * i.e. one line of a .fcp checkpoint (src/World.cpp:4234-4280). */
int ffh_blob_get_rng_state(ffh_blob *b, unsigned long long *state6);
int ffh_blob_set_rng_state(ffh_blob *b, const unsigned long long *state6);

/* Iterations of the last CG solve and its final ||r||^2/||f||^2 (for a batch: the largest
 * over its blobs). */
int ffh_blob_last_cg(ffh_blob *b, int *iterations, double *residual2);

/* Sizes: linear nodes, elements, 3x3 blocks of the matrix pattern, padded SpMV slots. */
int ffh_blob_sizes(ffh_blob *b, int *num_linear, int *num_elements, long long *num_blocks,
                   long long *num_slots);

/* The HIP stream (hipStream_t) all work of the blob is issued on. */
void *ffh_blob_stream(ffh_blob *b);

/* Wait for all work issued on the blob's stream. */
int ffh_blob_synchronize(ffh_blob *b);

/*
 * Per-kernel timing with HIP events on the blob's stream.  When enabled, every launch is
 * bracketed by events; ffh_blob_kernel_times returns, for each kernel id, the summed
 * milliseconds and launch counts since the last reset.  Kernel ids:
 *   0 element   1 assemble   2 rhs+CG init   3 CG spmv (p update, q = K p, p.q)
 *   4 CG update (x, r, residual)   5 integrate
 *   learned preconditioner: 6 couplings A_ij   7 embedding   8 message (A h)
 *   9 dense layer (MFMA)   10 output layer
 *   11 per-blob totals of the CG reductions (blobs of more than 3072 64-row slices)
 *   12 surface-surface interaction forces   13 their face geometry and lookup grid
 *   14 periodic box wrap
 */
#define FFH_NUM_KERNELS 15
int ffh_blob_set_profiling(ffh_blob *b, int enable);
int ffh_blob_kernel_times(ffh_blob *b, double *ms, long long *counts);

/* Strain energy of Blob::make_measurements (src/Blob.cpp:1424-1432) from the last element
 * update, in meso energy units. */
int ffh_blob_strain_energy(ffh_blob *b, double *energy);

/*
 * Learned preconditioner.  No counterpart in the reference, whose massless solver is
 * Jacobi-preconditioned CG only (src/NoMassCGSolver.cpp:165-197, :278-292); these calls extend
 * the Solver behind the same solve().  A mesh message-passing network (features: the
 * Jacobi-scaled residual normalised per blob and the node's drag share; embedding 4 -> 32; three
 * layers h' = relu(W1 h + W2 A h + b) with A the trace-normalised blocks of the current matrix;
 * output 32 -> 3) adds a correction to the Jacobi step: z = D^-1 r + s_b * net.  Its dense layers
 * run on bf16 MFMA with fp32 accumulation.
 *
 * With FFH_PRECOND_LEARNED the solve is flexible CG (Polak-Ribiere beta) for at most
 * fallback_iterations iterations (0: the iteration count of the last plain solve); blobs not
 * converged by then restart Jacobi-PCG from their current iterate and run to the same test
 * ||r||^2 / ||f||^2 < eps^2.  Until a plain solve has been made, solves stay plain.
 *
 * Training data: ffh_mpnn_harvest(b, S) records, for the next plain solves, the residual r_k of
 * every iteration and its exact correction x* - x_k (device memory, up to S snapshots of all
 * blobs); ffh_mpnn_train then runs Adam over the snapshots (one snapshot per step, seeded
 * order) with the matrix current at training time, and returns the mean loss per epoch.  The
 * loss of a snapshot is the mean over blobs of |net - t|^2 / |(x* - x_k) / s_b|^2 with
 * t = (x* - x_k - D^-1 r_k) / s_b.
 */
#define FFH_PRECOND_JACOBI 0
#define FFH_PRECOND_LEARNED 1
int ffh_mpnn_num_params(void);
int ffh_mpnn_init(ffh_blob *b, unsigned long long seed);
int ffh_mpnn_set_weights(ffh_blob *b, const float *weights);
int ffh_mpnn_get_weights(ffh_blob *b, float *weights);
int ffh_blob_set_preconditioner(ffh_blob *b, int kind, int fallback_iterations);
/* z = the learned preconditioner applied to the host residual r ([3*N_t]) on the current matrix. */
int ffh_mpnn_apply(ffh_blob *b, const double *r, double *z);
int ffh_mpnn_harvest(ffh_blob *b, int max_snapshots);
int ffh_mpnn_harvest_count(ffh_blob *b, int *snapshots, long long *pairs);
int ffh_mpnn_train(ffh_blob *b, int epochs, float learning_rate, unsigned long long seed, float *loss_per_epoch);
/* Loss and gradient (ffh_mpnn_num_params floats) of one (residual, correction) pair, [3*N_t]
 * each, without a weight update. */
int ffh_mpnn_loss_grad(ffh_blob *b, const double *r, const double *correction, double *loss, float *grad);

#ifdef __cplusplus
}
#endif
#endif
