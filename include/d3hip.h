/*
 * d3hip.h -- C ABI of libd3hip.so, the MI355X implementation of the DFT-D3
 * energy / gradient / virial path.
 *
 * Part 1 mirrors the subset of the simple-dftd3 v1.6.0 C API that its Python
 * layer binds for this path (include/s-dftd3.h of the reference): the same
 * symbol names, argument meaning, NULL conventions and error strings, with
 * host pointers.  A program or binding written against s-dftd3.h for these
 * entry points links against libd3hip.so unchanged.
 *
 * Part 2 adds entry points that take device pointers (e.g. PyTorch-ROCm
 * tensors) and a HIP stream, for single systems and batches of systems.
 *
 * Units: bohr and Hartree.  positions [n][3] row-major, lattice [3][3] with
 * rows = lattice vectors, gradient [n][3], sigma [3][3] in the order the
 * reference returns it (sigma(a,b) at [b][a]; the array is symmetric).
 */
#ifndef D3HIP_H
#define D3HIP_H

#ifdef __cplusplus
#define D3HIP_API extern "C"
#else
#include <stdbool.h>
#define D3HIP_API extern
#endif

typedef struct _dftd3_error* dftd3_error;
typedef struct _dftd3_structure* dftd3_structure;
typedef struct _dftd3_model* dftd3_model;
typedef struct _dftd3_param* dftd3_param;

/* s-dftd3.h:56-64: type-generic release of a handle (C11 _Generic), for the
 * handle types of this library (the reference's gcp handle is not part of it). */
#ifndef __cplusplus
#define dftd3_delete(ptr) _Generic((ptr), \
                       dftd3_error: dftd3_delete_error, \
                   dftd3_structure: dftd3_delete_structure, \
                       dftd3_model: dftd3_delete_model, \
                       dftd3_param: dftd3_delete_param \
                                  )(&ptr)
#endif

/* ------------------------------------------------------------------ Part 1 */

/* s-dftd3.h:71-72 dftd3_get_version: major*10000 + minor*100 + patch of the
 * API this library mirrors (1.6.0). */
D3HIP_API int dftd3_get_version(void);

/* s-dftd3.h:74-78 dftd3_has_feature (api.f90:150-162, feature.F90:37-52):
 * "mpi" is the only optional feature and this library has none of them. */
D3HIP_API bool dftd3_has_feature(char* feature);

/* s-dftd3.h:85-100 error handle (api.f90:176-238). */
D3HIP_API dftd3_error dftd3_new_error(void);
D3HIP_API int dftd3_check_error(dftd3_error error);
D3HIP_API void dftd3_get_error(dftd3_error error, char* buffer, const int* buffersize);
D3HIP_API void dftd3_delete_error(dftd3_error* error);

/* s-dftd3.h:107-113 dftd3_new_structure (api.f90:243-280): positions are
 * wrapped into the central cell; an error "Too close interatomic distances
 * found" is set for any pair closer than 1e-9 bohr.  lattice and periodic may
 * be NULL; a lattice without periodicity is periodic in three dimensions. */
D3HIP_API dftd3_structure dftd3_new_structure(dftd3_error error, const int natoms, const int* numbers,
                                              const double* positions, const double* lattice,
                                              const bool* periodic);

/* s-dftd3.h:116-117 (api.f90:284-296) */
D3HIP_API void dftd3_delete_structure(dftd3_structure* mol);

/* s-dftd3.h:120-124 dftd3_update_structure (api.f90:300-335); lattice may be NULL. */
D3HIP_API void dftd3_update_structure(dftd3_error error, dftd3_structure mol, const double* positions,
                                      const double* lattice);

/* s-dftd3.h:131-133 dftd3_new_d3_model (api.f90:339-364). */
D3HIP_API dftd3_model dftd3_new_d3_model(dftd3_error error, dftd3_structure mol);

/* s-dftd3.h:138-143 (api.f90:369-390), widths set to zero. */
D3HIP_API void dftd3_set_model_realspace_cutoff(dftd3_error error, dftd3_model model, double disp2, double disp3,
                                                double cn);

/* s-dftd3.h:146-150 dftd3_set_model_ghost_index: ghost atoms (1-based
 * indices) contribute no dispersion (model.f90:166-174). */
D3HIP_API void dftd3_set_model_ghost_index(dftd3_error error, dftd3_model model, const int* ghost, int nidx);

/* s-dftd3.h:153-160 (api.f90:393-416) */
D3HIP_API void dftd3_set_model_realspace_cutoff_smooth(dftd3_error error, dftd3_model model, double disp2,
                                                       double disp3, double cn, double width2, double width3);

/* s-dftd3.h:172-184 dftd3_set_model_ewald (api.f90:419-453): a low-rank
 * expansion of the reference C6 (rank > 0 fixed, else the smallest rank
 * within tolerance, default 1e-4) for every C6 of the model, and under 3D
 * periodic boundary conditions the two-body term summed over the reciprocal
 * lattice within kcut (> 0, else from the damping radii): mesh < 0 sums the
 * vectors directly (fourier/ewald.f90), mesh = 0 uses a particle mesh derived
 * from the cutoff and mesh > 0 that many points per direction, rounded up to
 * a power of two (fourier/spme.f90). */
D3HIP_API void dftd3_set_model_ewald(dftd3_error error, dftd3_model model, int rank, double tolerance, double kcut,
                                     int mesh);

/* s-dftd3.h:185-196 dftd3_set_model_work_partition (api.f90:459-480,
 * partition.f90:57-148): evaluate only part `part` (zero based) of `nparts`
 * of the pair, triple and reciprocal-space loops; the structure dependent
 * quantities (CN, weights, C6) and the CN chain rule stay whole, so summing
 * the results of all parts gives the complete calculation.  Applies to
 * dftd3_get_dispersion and dftd3_get_dispersion_hessian.  An invalid
 * partition ("Invalid dispersion work partition") leaves the model serial. */
D3HIP_API void dftd3_set_model_work_partition(dftd3_error error, dftd3_model model, int part, int nparts);

/* s-dftd3.h:198-205 dftd3_set_model_mpi_comm (api.f90:484-507): this library
 * has no MPI backend; like a reference build without MPI it reports
 * "s-dftd3 was built without MPI support" (and leaves the model serial). */
D3HIP_API void dftd3_set_model_mpi_comm(dftd3_error error, dftd3_model model, int comm);

/* No counterpart in the reference (its d3_lowrank_c6%rank and %error):
 * the rank and largest relative error of the model's low-rank expansion
 * (0 and 0 without one); the reciprocal cutoff and the number of vectors
 * (one of each +-k pair and k = 0) of the last reciprocal-space evaluation.
 * Any output may be NULL. */
D3HIP_API void d3hip_get_model_lowrank(dftd3_error error, dftd3_model model, int* rank, double* max_error,
                                       double* kcut, long long* vectors);

/* s-dftd3.h:208-209 (api.f90) */
D3HIP_API void dftd3_delete_model(dftd3_model* disp);

/* s-dftd3.h:216-229 zero damping (api.f90:625-694, param.f90:805-1090). */
D3HIP_API dftd3_param dftd3_new_zero_damping(dftd3_error error, double s6, double s8, double s9, double rs6,
                                             double rs8, double alp);
D3HIP_API dftd3_param dftd3_load_zero_damping(dftd3_error error, char* method, bool atm);

/* s-dftd3.h:232-245 rational damping (api.f90:555-622, param.f90:302-800). */
D3HIP_API dftd3_param dftd3_new_rational_damping(dftd3_error error, double s6, double s8, double s9, double a1,
                                                 double a2, double alp);
D3HIP_API dftd3_param dftd3_load_rational_damping(dftd3_error error, char* method, bool atm);

/* s-dftd3.h:248-262 modified zero damping (api.f90:748-813, mzero.f90):
 * zero damping with the offset bet. */
D3HIP_API dftd3_param dftd3_new_mzero_damping(dftd3_error error, double s6, double s8, double s9, double rs6,
                                              double rs8, double alp, double bet);
D3HIP_API dftd3_param dftd3_load_mzero_damping(dftd3_error error, char* method, bool atm);

/* s-dftd3.h:265-278 modified rational damping (api.f90:694-744): the rational
 * damping function with the parameters of param.f90:1091-1136. */
D3HIP_API dftd3_param dftd3_new_mrational_damping(dftd3_error error, double s6, double s8, double s9, double a1,
                                                  double a2, double alp);
D3HIP_API dftd3_param dftd3_load_mrational_damping(dftd3_error error, char* method, bool atm);

/* s-dftd3.h:281-295 optimized power damping (api.f90:817-882,
 * optimizedpower.f90). */
D3HIP_API dftd3_param dftd3_new_optimizedpower_damping(dftd3_error error, double s6, double s8, double s9,
                                                       double a1, double a2, double alp, double bet);
D3HIP_API dftd3_param dftd3_load_optimizedpower_damping(dftd3_error error, char* method, bool atm);

/* s-dftd3.h:298-312 CSO damping (api.f90:886-951, cso.f90): C6-only with a
 * sigmoidal scaling. */
D3HIP_API dftd3_param dftd3_new_cso_damping(dftd3_error error, double s6, double s9, double a1, double a2,
                                            double a3, double a4, double alp);
D3HIP_API dftd3_param dftd3_load_cso_damping(dftd3_error error, char* method, bool atm);

/* s-dftd3.h:315-327 Z damping (api.f90:955-1018, z.f90): r0 = a1 / (id_i +
 * id_j) with the species identifiers of the structure (elements numbered by
 * first appearance), damping C6 r0.  The loader has no entries
 * (param.f90:1297-1326): every method name is an error. */
D3HIP_API dftd3_param dftd3_new_z_damping(dftd3_error error, double s6, double s8, double s9, double a1,
                                          double alp);
D3HIP_API dftd3_param dftd3_load_z_damping(dftd3_error error, char* method, bool atm);

/* s-dftd3.h:330-331 */
D3HIP_API void dftd3_delete_param(dftd3_param* param);

/* s-dftd3.h:380-388 dftd3_get_dispersion (api.f90:1038-1111): gradient and
 * sigma may be NULL; requesting either evaluates both derivatives. */
D3HIP_API void dftd3_get_dispersion(dftd3_error error, dftd3_structure mol, dftd3_model disp, dftd3_param param,
                                    double* energy, double* gradient, double* sigma);

/* s-dftd3.h:390-397 dftd3_get_pairwise_dispersion (api.f90:1180-1233,
 * disp.f90:403-457): pair_energy2[n][n] receives the additive (two-body) and
 * pair_energy3[n][n] the non-additive (three-body) energy of every pair of
 * atoms (host arrays; both matrices are symmetric). */
D3HIP_API void dftd3_get_pairwise_dispersion(dftd3_error error, dftd3_structure mol, dftd3_model disp,
                                             dftd3_param param, double* pair_energy2, double* pair_energy3);

/* s-dftd3.h:399-407 dftd3_get_dispersion_hessian (api.f90:1115-1176,
 * disp.f90:236-308): the energy and the analytic second derivatives
 * hessian[3n][3n] (host; symmetric, index 3*i + c for component c of atom
 * i), in Hartree / bohr^2.  Always evaluated in FP64. */
D3HIP_API void dftd3_get_dispersion_hessian(dftd3_error error, dftd3_structure mol, dftd3_model disp,
                                            dftd3_param param, double* energy, double* hessian);

/* ------------------------------------------------------------------ Part 2
 * No counterpart in the reference: device-pointer entry points. */

/* Arithmetic of the pair and triple terms: FP64 (default) or FP32 with FP64
 * accumulation of energies, gradients and virials. */
#define D3HIP_PRECISION_FP64 0
#define D3HIP_PRECISION_MIXED 1

D3HIP_API void d3hip_set_model_precision(dftd3_error error, dftd3_model model, int precision);

/* Evaluate a batch of independent systems in one pass.  Host arrays:
 *   natoms[nsys], numbers[sum natoms], lattices[nsys][3][3] (may be NULL
 *   when no system is periodic), periodic[nsys][3] (may be NULL),
 *   params[nsys][10] = {kind, s6, s8, s9, a1|rs6, a2|rs8, alp, bet, a3, a4}
 *   with kind 0 rational (also the modified rational parameters), 1 zero,
 *   2 modified zero, 3 optimized power, 4 CSO, 5 Z damping (rs6, rs8 in the
 *   a1, a2 places for zero and modified zero; unused entries are ignored),
 *   cutoffs[nsys][5] = {disp2, disp3, cn, width2, width3} (NULL selects the
 *   defaults 60, 40, 40, 0, 0).
 * Device arrays: positions[sum natoms][3] (not wrapped by this call),
 *   energy[nsys], gradient[sum natoms][3] (may be NULL), sigma[nsys][3][3]
 *   (may be NULL).  stream is a hipStream_t (NULL = default stream).
 * The call is asynchronous with respect to the host except for one
 * device-to-host read of the three-body list size. */
D3HIP_API void d3hip_get_dispersion_batch(dftd3_error error, dftd3_model workspace, int nsys, const int* natoms,
                                          const int* numbers, const double* lattices, const bool* periodic,
                                          const double* params, const double* cutoffs, int precision,
                                          const double* positions, double* energy, double* gradient,
                                          double* sigma, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): d3hip_get_dispersion_batch without a workspace and
 * without any allocation.  Every work buffer comes from a fixed work area of
 * D3HIP_NOALLOC_BYTES per GPU, a static device array of the library (held
 * until the process ends); the metadata is copied from pageable host memory.
 * A batch that needs more than the area (about 72 n^2 bytes for an FP64
 * molecule of n atoms with the three-body term, 44 n^2 mixed: n up to some
 * 950 / 1200) is rejected with an error before any output is written.  No
 * low-rank model.  Calls that overlap on one GPU race: they share its area,
 * which keeps the last call's contents.  Same arguments and results as
 * d3hip_get_dispersion_batch otherwise (including the read-back of the
 * three-body list size). */
#define D3HIP_NOALLOC_BYTES (64u << 20)
D3HIP_API void d3hip_get_dispersion_batch_noalloc(dftd3_error error, int nsys, const int* natoms,
                                                  const int* numbers, const double* lattices, const bool* periodic,
                                                  const double* params, const double* cutoffs, int precision,
                                                  const double* positions, double* energy, double* gradient,
                                                  double* sigma, void* stream);

/* Pairwise-resolved energies of a batch (device arrays; same batch
 * description as d3hip_get_dispersion_batch): pair_energy2 and pair_energy3
 * hold per system an n x n block in atom order, the blocks one after another
 * (system s at offset sum of n^2 of the systems before it).  Asynchronous
 * except for one device-to-host read of the three-body list size. */
D3HIP_API void d3hip_get_pairwise_dispersion_batch(dftd3_error error, dftd3_model workspace, int nsys,
                                                   const int* natoms, const int* numbers, const double* lattices,
                                                   const bool* periodic, const double* params,
                                                   const double* cutoffs, int precision, const double* positions,
                                                   double* pair_energy2, double* pair_energy3, void* stream);

/* Energies and analytic Hessians of a batch, FP64 (device arrays; same batch
 * description as d3hip_get_dispersion_batch): energy[nsys] (may be NULL) and
 * hessian holding per system a 3n x 3n block, the blocks one after another
 * (system s at offset 9 times the sum of n^2 of the systems before it).
 * Asynchronous except for one device-to-host read of the three-body list
 * size. */
D3HIP_API void d3hip_get_dispersion_hessian_batch(dftd3_error error, dftd3_model workspace, int nsys,
                                                  const int* natoms, const int* numbers, const double* lattices,
                                                  const bool* periodic, const double* params, const double* cutoffs,
                                                  const double* positions, double* energy, double* hessian,
                                                  void* stream);

/* A model without a structure, usable only as the workspace (device buffers
 * and stream) of d3hip_get_dispersion_batch. */
D3HIP_API dftd3_model d3hip_new_workspace(dftd3_error error);

/* Count the work units the reference's loops visit, per system, on the GPU:
 * counts[nsys][3] (host) = {CN terms (ncoord.f90:118-127), two-body pair
 * terms (rational.f90:450-464), ATM triple terms (atm.f90:706-742)}.  Same
 * batch description as d3hip_get_dispersion_batch; triples are counted even
 * for s9 = 0.  Synchronizes the stream. */
D3HIP_API void d3hip_count_terms_batch(dftd3_error error, dftd3_model workspace, int nsys, const int* natoms,
                                       const int* numbers, const double* lattices, const bool* periodic,
                                       const double* params, const double* cutoffs, const double* positions,
                                       long long* counts, void* stream);

/* Time every kernel launch of the model / workspace with HIP events on its
 * stream (enable != 0; resets the totals).  Timing synchronizes each call. */
D3HIP_API void d3hip_set_kernel_timing(dftd3_error error, dftd3_model model, int enable);

/* Total milliseconds and launches of one kernel group since timing was
 * enabled; name is one of pack, order, cn, weights, c6, disp2, atm_list, atm,
 * cn_chain, reduce, hessian. */
D3HIP_API void d3hip_get_kernel_time(dftd3_error error, dftd3_model model, const char* name, double* total_ms,
                                     long long* calls);

/* ------------------------------------------------------------------ plans
 * A batch planned once and evaluated many times with new coordinates of the
 * same systems (no counterpart in the reference).  The metadata upload,
 * buffer sizes and three-body launch shape are fixed by the plan; an
 * evaluation is a fixed sequence of kernel launches on the stream without
 * host synchronisation or device allocation, and can be recorded into a HIP
 * graph (d3hip_plan_capture) and replayed (d3hip_plan_replay).
 * Three-body lists: molecules get the exact bound; periodic systems get 1.5x
 * the list size at the planning coordinates (at least 64 more slots per
 * atom of a system with a three-body term).  An evaluation whose lists do
 * not fit adds no three-body terms, writes NaN into the energy, gradient and
 * virial of every system with a three-body term, and sets the flag that
 * d3hip_plan_overflowed reports. */
typedef struct _d3hip_plan* d3hip_plan;

/* Same batch description as d3hip_get_dispersion_batch; grad != 0 plans
 * energy + gradient + virial, else the energy only; positions (device) are
 * the planning coordinates.  Synchronizes the stream. */
D3HIP_API d3hip_plan d3hip_new_plan(dftd3_error error, int nsys, const int* natoms, const int* numbers,
                                    const double* lattices, const bool* periodic, const double* params,
                                    const double* cutoffs, int precision, int grad, const double* positions,
                                    void* stream);
D3HIP_API void d3hip_delete_plan(d3hip_plan* plan);

/* One evaluation; device outputs as in d3hip_get_dispersion_batch (gradient
 * and sigma must be given exactly when the plan has grad != 0). */
D3HIP_API void d3hip_plan_execute(dftd3_error error, d3hip_plan plan, const double* positions, double* energy,
                                  double* gradient, double* sigma, void* stream);

/* Candidate launch shapes of the three-body kernel: wavefronts per block
 * (1, 2 or 4) and persistent blocks per compute unit (0 = one block per work
 * item).  A new plan starts with the shape of a call without a plan: four
 * wavefronts and one block per item below 8,192 work items; above, two
 * wavefronts and 8 blocks per CU, or one and 16 for mixed-precision lists
 * with lattice images. */
D3HIP_API int d3hip_plan_num_shapes(void);
D3HIP_API void d3hip_plan_candidate_shape(int index, int* waves, int* blocks_per_cu);
D3HIP_API void d3hip_plan_get_shape(dftd3_error error, d3hip_plan plan, int* waves, int* blocks_per_cu);
D3HIP_API void d3hip_plan_set_shape(dftd3_error error, d3hip_plan plan, int waves, int blocks_per_cu);

/* Time every candidate shape on this batch (one warm-up and reps timed
 * evaluations each) and keep the fastest; returns its index, -1 on error.
 * times_ms[d3hip_plan_num_shapes()] (host, may be NULL) receives the mean
 * milliseconds per evaluation of each.  Synchronizes the stream. */
D3HIP_API int d3hip_plan_autotune(dftd3_error error, d3hip_plan plan, const double* positions, double* energy,
                                  double* gradient, double* sigma, int reps, float* times_ms, void* stream);

/* Record one evaluation with these device pointers into a HIP graph (the
 * stream must not be capturing already); replay launches it on a stream.
 * New coordinates go into the same positions buffer before a replay. */
D3HIP_API void d3hip_plan_capture(dftd3_error error, d3hip_plan plan, const double* positions, double* energy,
                                  double* gradient, double* sigma, void* stream);
D3HIP_API void d3hip_plan_replay(dftd3_error error, d3hip_plan plan, void* stream);

/* 1 when an evaluation since the plan was made had more three-body slots than
 * the plan holds, 0 otherwise, -1 on error.  Synchronizes the stream. */
D3HIP_API int d3hip_plan_overflowed(dftd3_error error, d3hip_plan plan, void* stream);

#endif /* D3HIP_H */
