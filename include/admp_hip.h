/*
 * admp_hip.h -- C ABI of the MI355X (gfx950) ADMP multipolar PME library.
 *
 * The library evaluates DMFF's ADMP multipolar PME energy, the induced-dipole
 * solve, dispersion PME and the Slater short-range term, plus their gradients,
 * with hand-written HIP kernels.  It is the drop-in boundary for the reference's
 * ADMPPmeForce / ADMPDispPmeForce backends:
 *
 *   reference                                    this ABI
 *   dmff/admp/pme.py:36-114   ADMPPmeForce(...)    admp_ctx_create + admp_set_topology
 *   dmff/common/nblist.py:27-55 NeighborList.allocate  admp_build_neighbors + admp_get_pairs
 *   dmff/admp/pme.py:186-232  get_energy (pol)     admp_set_pairs + admp_compute
 *   dmff/admp/pme.py:120-141  get_energy (non-pol) admp_set_pairs + admp_compute
 *   dmff/admp/pme.py:339      get_forces = value_and_grad   admp_compute(want_grad=1)
 *   dmff/admp/pme.py:342-423  optimize_Uind        admp_compute (solver options)
 *   dmff/admp/disp_pme.py:13-77 ADMPDispPmeForce.get_energy  admp_disp_compute
 *   dmff/admp/pairwise.py:47-95,156-169 slater_sr_kernel       admp_slater_compute
 *   dmff/admp/pairwise.py:47-169 generate_pairwise_interaction   admp_pairwise_compute
 *   dmff/admp/pme.py:298-311  Q_local_to_global    admp_local_to_global
 *   dmff/admp/pme.py:236-296  generate_esp         admp_local_to_global + admp_esp
 *   openmm_dmff_plugin CudaDMFFKernels.cpp:200-220 addForces    admp_add_forces_fixed
 *
 * Conventions
 *   - Units are the reference backend's: Angstrom, e, kJ/mol (DIELECTRIC =
 *     1389.35455846).  Conversion from nm happens in the Python host layer.
 *   - Every array argument of the compute entry points is a DEVICE pointer owned
 *     by the caller, in row-major order, float64 unless stated.  Host pointers
 *     are marked (host).
 *   - Boxes are 3x3 with the lattice vectors in rows (reference convention).
 *   - A context owns all scratch (neighbour lists, grids, FFT twiddles, solver
 *     vectors).  Calls on one context are asynchronous on its stream except where
 *     a function documents a host synchronisation.  One context per thread.
 *   - Every function returns an admp_status; 0 is success.  No C++ exception
 *     crosses the ABI; admp_last_error() describes the last failure.
 *   - A batch of n_frames configurations of the same topology can be evaluated
 *     in one call: per-frame arrays are stacked on a leading frame axis.
 */
#ifndef ADMP_HIP_H
#define ADMP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ADMP_OK = 0,
    ADMP_ERR_ARG = 1,          /* bad argument / shape                         */
    ADMP_ERR_CAPACITY = 2,     /* neighbour capacity exceeded (count reported)  */
    ADMP_ERR_HIP = 3,          /* HIP runtime error                             */
    ADMP_ERR_STATE = 4,        /* call order (e.g. compute before topology)     */
    ADMP_WARN_NOT_CONVERGED = 5/* solver hit max iterations; results are valid  */
} admp_status;

typedef enum { ADMP_FP64 = 0, ADMP_MIXED = 1 } admp_precision;

/* Induced-dipole solver.
 *  ADMP_SOLVER_REFERENCE: the reference's Jacobi update U <- U - dE/dU*pol/DIELECTRIC,
 *      stop when max|dE/dU| over sites with pol > 0.001 A^3 is below tol (reference
 *      POL_CONV = 1.0) or after maxiter (MAX_N_POL = 30) sweeps (pme.py:380-401).
 *  ADMP_SOLVER_STEPS: exactly `steps` Jacobi sweeps (pme.py:404-422, steps_pol).
 *  ADMP_SOLVER_PCG: preconditioned conjugate gradients to the same fixed point,
 *      stop on the same max-norm criterion (used for tight tolerances). */
typedef enum { ADMP_SOLVER_REFERENCE = 0, ADMP_SOLVER_STEPS = 1, ADMP_SOLVER_PCG = 2 } admp_solver_kind;
typedef enum { ADMP_PRECOND_CLUSTERS = 0, ADMP_PRECOND_DIAGONAL = 1, ADMP_PRECOND_PAIRS = 2 } admp_precond;

typedef struct {
    int32_t n_atoms;
    int32_t n_frames;          /* batch size (>= 1)                           */
    int32_t lmax;              /* 0, 1 or 2                                   */
    int32_t lpol;              /* induced dipoles on/off                      */
    int32_t lpme;              /* 0: kappa = 0 and no reciprocal/self (pme.py:83-88) */
    int32_t precision;         /* admp_precision                              */
    double  kappa;             /* 1/A (from the ethresh rule, pme.py:460-466) */
    int32_t K1, K2, K3;        /* PME grid                                    */
    int32_t pme_order;         /* must be 6 (recip.py:13)                     */
    double  rc;                /* neighbour cutoff, A                         */
    int32_t max_neighbors;     /* per-atom full-list capacity (0 = estimate)  */
    /* dispersion PME (0 grid = disabled) */
    double  disp_kappa;
    int32_t dK1, dK2, dK3;
    int32_t disp_pmax;         /* 6, 8 or 10                                  */
} admp_config;

typedef struct {
    int32_t kind;              /* admp_solver_kind                            */
    int32_t maxiter;
    int32_t steps;
    double  tol;               /* max |dE/dU| threshold, kJ/mol/(e*A)         */
    int32_t warm_start;        /* 1: use U_inout as the initial guess         */
    int32_t fp32_products;     /* PCG on a MIXED context: 1 = reciprocal products in
                                  FP32 until the true residual fell 1000x (or 32
                                  iterations), then a restart with FP64 products;
                                  0 = FP64 products throughout (default)      */
    int32_t precond;           /* PCG preconditioner: ADMP_PRECOND_CLUSTERS
                                  (default: exact blocks of up to 4 sites, a
                                  pair of mutually nearest coupled sites within
                                  2 A with the sites whose nearest coupled
                                  neighbour is in the pair; diagonal elsewhere),
                                  ADMP_PRECOND_PAIRS (the pairs alone) or
                                  ADMP_PRECOND_DIAGONAL (alpha / DIELECTRIC) */
} admp_solver_opts;

typedef struct {
    /* filled on return of admp_compute (host-visible; admp_compute synchronises) */
    int32_t n_iter[64];        /* per frame (first 64 frames)                 */
    int32_t converged[64];
    double  residual[64];      /* final max |dE/dU| over polarisable sites    */
    double  e_real, e_recip, e_self, e_pol;   /* frame-0 components         */
} admp_stats;

typedef struct admp_ctx admp_ctx;

const char* admp_version(void);

/* ADMPPmeForce.__init__ (pme.py:43-114): allocate grids, FFT tables, scratch. */
admp_status admp_ctx_create(const admp_config* cfg, admp_ctx** out);
admp_status admp_ctx_destroy(admp_ctx* ctx);
const char* admp_last_error(const admp_ctx* ctx);
/* stream: a hipStream_t (NULL = the legacy default stream). */
admp_status admp_set_stream(admp_ctx* ctx, void* stream);

/* No counterpart in the reference (a serving and MD option).  caller_mode 1:
 * the context works directly on the stream last given to admp_set_stream
 * (when it is not the null stream) instead of on its own stream ordered by
 * events against it.  For callers that drive many contexts from a few worker
 * streams: HIP maps streams onto a few hardware queues, so hundreds of
 * contexts with streams of their own share queues and wait on each other's
 * unrelated work.  A change of stream is ordered after the context's last
 * work (an event recorded by every entry point; in caller mode carried over to
 * the context's own stream, so that the wait never touches a stream another
 * thread may be capturing); captured solver graphs are replayed on whichever
 * stream is active.  Default 0. */
admp_status admp_set_stream_mode(admp_ctx* ctx, int32_t caller_mode);

/* Local frames (axis_type (N,), axis_idx (N,3) [z,x,y], -1 = none; spatial.py:44-156)
 * and the covalent exclusions (i, j, nbond) for nbond in 1..6
 * (dmff/api/topology.py:391-418, each unordered pair once).  All (host). */
admp_status admp_set_topology(admp_ctx* ctx, const int32_t* axis_type, const int32_t* axis_idx,
                              int64_t n_excl, const int32_t* excl_i, const int32_t* excl_j,
                              const int32_t* excl_nbond);

/* Reference-format pair list (Np,3) [i, j, nbond] of frame `frame`; rows with
 * i >= j (padding) are ignored (dmff/utils.py:65-84).  pairs_is_int64 selects
 * the element type.  No distance cutoff is applied to listed pairs (pme.py:1052-1173).
 * Synchronises the stream once (capacity check). */
admp_status admp_set_pairs(admp_ctx* ctx, int32_t frame, const void* pairs, int64_t n_pairs,
                           int32_t pairs_is_int64);

/* Cell-list neighbour build on the GPU for all frames: the pair set
 * {i<j : min-image |r_i - r_j| < rc} (nblist.py:27-55), distances in FP64 with
 * a fixed, unfused operation order.  pos (F,N,3), box (F,3,3).  The first
 * build reads the boxes back to size the cell grid and synchronises; later
 * builds take that grid over without waiting (a box it no longer fits is
 * rebuilt on a fresh grid by the next call that reads the lists, which also
 * reports a bad box or bad positions). */
admp_status admp_build_neighbors(admp_ctx* ctx, const double* pos, const double* box);

/* Write the i<j pairs of `frame` in the reference format (int32 (n,3) [i,j,nbond])
 * to device buffer `out` with room for `cap` rows; *n_out receives the count.
 * Synchronises. */
admp_status admp_get_pairs(admp_ctx* ctx, int32_t frame, int32_t* out, int64_t cap, int64_t* n_out);

/* Energy and (optionally) all gradients for every frame.
 *   pos (F,N,3) A, box (F,3,3) A, Q_local (N,9) harmonic per atom, pol (N,) A^3,
 *   thole (N,), m/p/d scales (host, n_scales each; index nbond-1 with
 *   Python wrap for nbond = 0, pme.py:1111-1125).
 *   U_inout (F,N,3) e*A global Cartesian: warm start in, induced dipoles out.
 *   energy (F,) out.  With want_grad: dpos (F,N,3), dbox (F,3,3), dQ_local (F,N,9),
 *   dpol (F,N), dthole (F,N) out (any may be NULL to skip writing it).
 *   Gradients hold U fixed (Hellmann-Feynman, pme.py:364-371).
 *   With stats: synchronises the stream once at the end (solver statistics,
 *   energy components).  stats NULL: returns as soon as the work is queued
 *   when the call has already waited on the device after its geometry check
 *   was copied back (a polarisable CG solve does, for its convergence reads);
 *   the outputs are ordered on the stream as always. */
admp_status admp_compute(admp_ctx* ctx, const double* pos, const double* box,
                         const double* Q_local, const double* pol, const double* thole,
                         const double* mScales, const double* pScales, const double* dScales,
                         int32_t n_scales, double* U_inout, const admp_solver_opts* opts,
                         int32_t want_grad, double* energy, double* dpos, double* dbox,
                         double* dQ_local, double* dpol, double* dthole, admp_stats* stats);

/* Dispersion PME (disp_pme.py:80-130): backend energy (positive convention; the
 * generator negates it) for c_list (N,3) = sqrt(C6, C8, C10) per atom.
 * Uses the context's neighbour list.  Gradients: dpos, dbox, dc (N,3). */
admp_status admp_disp_compute(admp_ctx* ctx, const double* pos, const double* box,
                              const double* c_list, const double* mScales, int32_t n_scales,
                              int32_t want_grad, double* energy, double* dpos, double* dbox,
                              double* dc);

/* Slater-ISA short range (pairwise.py:156-169): sum m*Ai*Aj*(1/3 (Br)^2+Br+1)e^-Br,
 * B = sqrt(Bi*Bj).  a (N,), b (N,) in kJ/mol and 1/A.  Gradients dpos, dbox, da, db. */
admp_status admp_slater_compute(admp_ctx* ctx, const double* pos, const double* box,
                                const double* a, const double* b, const double* mScales,
                                int32_t n_scales, int32_t want_grad, double* energy,
                                double* dpos, double* dbox, double* da, double* db);

/* Q_local_to_global (dmff/admp/pme.py:298-311) for every frame: local harmonic
 * multipoles Q_local (N, (lmax+1)^2) per site, rotated by the local frames of
 * pos (F,N,3) / box (F,3,3), to global Cartesian Q_global (F,N,10)
 * [q, mu_x, mu_y, mu_z, Theta_xx, Theta_yy, Theta_zz, Theta_xy, Theta_xz, Theta_yz]
 * in the reference's C1_h2c / C2_h2c convention.  Synchronises once. */
admp_status admp_local_to_global(admp_ctx* ctx, const double* pos, const double* box, const double* Q_local,
                                 double* Q_global);

/* generate_esp's esp_kernel (dmff/admp/pme.py:236-296): the electrostatic
 * potential (kJ/mol/e) at n_grid probe points grids (n_grid,3) from n_sites
 * global multipoles Q_global (n_sites,10, as admp_local_to_global writes them)
 * at pos (n_sites,3) and induced dipoles U (n_sites,3, e nm; NULL = none);
 * positions and probes in nm, no periodic images.  esp (n_grid,) out.  Device
 * pointers; needs no context; asynchronous on stream. */
admp_status admp_esp(int32_t n_sites, const double* pos, const double* Q_global, const double* U, int32_t n_grid,
                     const double* grids, double* esp, void* stream);

/* MD integration: CudaCalcDMFFForceKernel::execute's force hand-off
 * (backend/openmm_dmff_plugin/platforms/cuda/src/CudaDMFFKernels.cpp:200-220,
 * kernels/DMFFForce.cu:1-10).  Adds force = -scale * dE/dpos to OpenMM-style
 * fixed-point force buffers (long long, 2^32 units per kJ/mol/nm; layout
 * [x[padded], y[padded], z[padded]]): slot a receives the atom atom_index[a]
 * (NULL = identity).  scale converts the library's kJ/mol/A to the buffer's
 * unit (10 for kJ/mol/nm).  dE_dpos (n_atoms,3), atom_index, force_buffers:
 * device pointers; asynchronous on stream; needs no context. */
admp_status admp_add_forces_fixed(const double* dE_dpos, double scale, const int32_t* atom_index,
                                  int64_t* force_buffers, int32_t n_atoms, int32_t padded_n_atoms, void* stream);

/* Short-range pair kernels of generate_pairwise_interaction (dmff/admp/pairwise.py:47-95)
 * and their per-site parameters, in the kernel's argument order:
 *   ADMP_SR_SLATER       slater_sr_kernel (pairwise.py:156-169)           (A, B)
 *   ADMP_SR_TT_QQ_C6     TT_damping_qq_c6_kernel (pairwise.py:98-117)     (A, B, Q, C)
 *   ADMP_SR_TT_QQ        TT_damping_qq_kernel (pairwise.py:120-128)       (B, Q)
 *   ADMP_SR_SLATER_DAMP  slater_disp_damping_kernel (pairwise.py:131-153) (B, C6, C8, C10)
 * Parameters are in the kernels' own units (the generators' conversions,
 * dmff/generators/admp.py:190-199,508-509,640-643,772-773, happen in the host layer). */
typedef enum {
    ADMP_SR_SLATER = 1, ADMP_SR_TT_QQ_C6 = 2, ADMP_SR_TT_QQ = 3, ADMP_SR_SLATER_DAMP = 4
} admp_sr_kind;

/* pair_int(positions, box, pairs, mScales, *atomic_params) of
 * generate_pairwise_interaction (pairwise.py:47-95) for every frame, over the
 * context's pair lists (admp_set_pairs / admp_build_neighbors): energy (F,) and,
 * with want_grad, dpos (F,N,3), dbox (F,3,3) and dprm (F,N,n_prm), the gradient
 * with respect to each site's parameters.  prm (N, n_prm) device, n_prm the
 * kind's parameter count.  Synchronises the stream once. */
admp_status admp_pairwise_compute(admp_ctx* ctx, int32_t kind, const double* pos, const double* box,
                                  const double* prm, int32_t n_prm, const double* mScales, int32_t n_scales,
                                  int32_t want_grad, double* energy, double* dpos, double* dbox, double* dprm);

/* No counterpart in the reference: a variant of admp_pairwise_compute for callers
 * that cannot allocate (code under HIP graph capture, tight serving or MD loops).
 * It needs no context and takes the reference-format pair list (n_pairs, 3)
 * [i, j, nbond] (int32, or int64 with pairs_is_int64) of one configuration
 * directly, as pair_int does (dmff/admp/pairwise.py:47-95); rows that are not
 * live pairs (i >= j, (N, N) padding, indices outside [0, n_atoms)) contribute
 * nothing.  pos (n_atoms,3), box (3,3), prm (n_atoms, n_prm), outputs energy (1),
 * dpos (n_atoms,3), dbox (3,3), dprm (n_atoms, n_prm): device pointers.
 * precision: ADMP_FP64 or ADMP_MIXED (FP32 pair arithmetic, FP64 sums).
 * mScales: host.  stream: a hipStream_t (NULL = the legacy default stream).
 * It allocates nothing and does not synchronise: four kernel launches on
 * `stream`.  Its scratch (box record and per-block energy/virial partials) is
 * one static device array of 82,432 bytes per GPU, held until the process
 * ends; its size does not depend on the input.  Calls that overlap on one GPU
 * race on it.  Forces and parameter gradients are summed with atomic adds
 * (rounding order not fixed); energy and dbox in a fixed order.  An unusable
 * box (non-finite, or a height below 0.5 A) yields NaN energy and dbox.
 * Argument errors return ADMP_ERR_ARG before any launch; admp_last_error(NULL)
 * describes the last such error of the calling thread. */
admp_status admp_pairwise_compute_noalloc(int32_t kind, int32_t precision, int32_t n_atoms, const double* pos,
                                          const double* box, const void* pairs, int64_t n_pairs,
                                          int32_t pairs_is_int64, const double* prm, int32_t n_prm,
                                          const double* mScales, int32_t n_scales, int32_t want_grad,
                                          double* energy, double* dpos, double* dbox, double* dprm, void* stream);

/* Instrumentation (no counterpart in the reference): per-kernel-family timing with
 * HIP events recorded around the library's launches on the context stream.
 * admp_profile switches recording on/off (reset != 0 clears the totals);
 * admp_profile_get synchronises the stream and returns the accumulated milliseconds
 * and launch count of family `key` (0 real_final, 1 real_permfield, 2 real_matvec,
 * 3 spline_prep, 4 spread, 5 fft, 6 influence, 7 gather, 8 nbr_build, 9 frames,
 * 10 pair_short, 11 solver_vec, 12 solver_graph, 13 plane_matvec) and its name. */
admp_status admp_profile(admp_ctx* ctx, int32_t enable, int32_t reset);
admp_status admp_profile_get(admp_ctx* ctx, int32_t key, double* total_ms, int64_t* count, const char** name);

/* Per-type parameter gradients: out[f][t][c] = sum over sites i with
 * type_map[i] == t of per_site[f][i][c], summed in a fixed order (results do not
 * depend on scheduling).  The adjoint of the per-type -> per-site expansion
 * Q_local[map_atomtype] (dmff/generators/admp.py:1372-1375), which jax.grad
 * forms in the reference.  per_site (F,N,d), type_map (N,), out (F,n_types,d):
 * device pointers; stream: a hipStream_t (NULL = the legacy default stream).
 * Needs no context. */
admp_status admp_type_reduce(const double* per_site, int32_t n_frames, int32_t n_sites, int32_t d,
                             const int32_t* type_map, int32_t n_types, double* out, void* stream);

/* Instrumentation (no counterpart in the reference): device bytes currently held
 * by the context (grids, lists, scratch).  Serving uses it to admit requests by
 * HBM footprint.  Host-only; no GPU work. */
admp_status admp_memory_bytes(const admp_ctx* ctx, int64_t* bytes);

#ifdef __cplusplus
}
#endif
#endif /* ADMP_HIP_H */
