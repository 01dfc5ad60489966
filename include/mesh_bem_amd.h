/*
 * mesh_bem_amd -- C ABI of the MI355X (gfx950) mesh-BEM library.
 *
 * Conventions (every entry point):
 *   - Array arguments are DEVICE pointers, scalars are passed by value.
 *   - prec selects the element type of every floating-point array of the call:
 *     MBA_F32 -> float / complex float, MBA_F64 -> double / complex double.
 *     Arrays documented as "double" are double regardless of prec.
 *   - Complex arrays are interleaved (re, im), i.e. the memory of a torch
 *     complex64 / complex128 tensor.  The reference's op-level kernels write
 *     separate real and imaginary planes that its Python wrapper recombines
 *     (src/mesh_BEM/kernel_functions.py:52-72); here the kernels write the
 *     combined complex array directly.
 *   - Matrices are row-major with an explicit leading dimension (elements).
 *   - Index arrays are int32 (as the reference's faces / face_edge_indices).
 *   - Work is enqueued on `stream` (a hipStream_t).  Calls return without
 *     waiting for the device, except mba_lu_factor and mba_lu_factor_profiled
 *     (to read the panel wait flag; mba_lu_factor_async does not wait) and the
 *     _profiled batched factorisation (to read its clock).  Calls that take a
 *     host job table copy it into pinned staging memory of the library (a
 *     ring of slots per device, each reused after its copy has completed), so
 *     the caller's table may be released on return; the device scratch and
 *     every buffer passed must stay allocated until the stream's work is
 *     done.
 *   - Return value: 0 (MBA_OK) or an MBA_ERR_* code.  Argument and size errors
 *     are detected on the host before any launch.  No C++ exception leaves an
 *     entry point (allocation failures return MBA_ERR_ALLOC).
 *   - Memory: the caller passes every output and workspace, except that
 *     mba_mom_zmat_faceedge_ref, mba_mom_rhs_faceedge_ref and
 *     mba_mom_rhs_faceedge_ns take their per-face records from the stream-
 *     ordered allocator (hipMallocAsync / hipFreeAsync on `stream`),
 *     mba_mom_zmat_faceedge_ref_noalloc uses a fixed device array, and the
 *     device measurement mba_measure_valu_rates allocates its own buffer.
 *   - State: calls on different streams are independent except through the
 *     buffers the caller passes and these process-wide defaults, each an
 *     atomic value set by its setter: the FP32 GEMM tile variant
 *     (mba_gemm_set_variant) and the split-FP16 tile (mba_gemm_split_f16_set_tile),
 *     which change only how a product is tiled and may take effect inside a
 *     running call; and the complex64 LU's trailing product, panel and
 *     solve kinds (mba_lu_set_update / _panel / _solve), read once at the
 *     start of each factorisation or solve.  mba_mom_zmat_faceedge_ref_noalloc
 *     shares one device array per GPU between all its calls.
 */
#ifndef MESH_BEM_AMD_H
#define MESH_BEM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *mba_stream_t;

enum {
    MBA_OK = 0,
    MBA_ERR_ARG = 1,     /* null pointer, bad enum, bad leading dimension */
    MBA_ERR_SIZE = 2,    /* size out of the supported range */
    MBA_ERR_MESH = 3,    /* malformed mesh (detected on the host side) */
    MBA_ERR_ALLOC = 4,   /* workspace too small */
    MBA_ERR_LAUNCH = 5,  /* kernel launch failed */
    MBA_ERR_NUMERIC = 6, /* zero pivot / non-finite value reported by a kernel */
    MBA_ERR_TIMEOUT = 7, /* a cooperative kernel's bounded wait expired */
    MBA_ERR_IO = 8       /* file missing or unreadable / unwritable (host mesh I/O) */
};

enum { MBA_F32 = 0, MBA_F64 = 1 };
/* Scalar field of the FP32 GMRES building blocks. */
enum { MBA_COMPLEX = 0, MBA_REAL = 1 };

/* Human-readable text of a status code. */
const char *mba_status_string(int status);
/* Library ABI version (major * 100 + minor). */
int mba_version(void);

/* ===================================================================== MoM
 * RWG basis on face f, local edge e (edge opposite vertex faces[f][e]):
 *   f(x) = s * l / (2 A) * (x - v_e),  s = +1 iff faces[f][(e+1)%3] < faces[f][(e+2)%3]
 * (src/mesh_BEM/cuda/mesh_MOM_3d.cu:154-242).  "Face-edge" index = 3 f + e. */

/* Face-edge impedance blocks, reference rule: replaces
 * create_MOM_Zmat_3d_forward (src/mesh_BEM/cuda/mesh_MOM_3d.cpp:52-77,
 * kernel mesh_MOM_3d.cu:332-496).  CFIE (EFIE + MFIE, eta = 1), test rule =
 * collapsed Gauss (gp_x, gw_x) of order nq_x, source rule (gp_y, gw_y) of
 * order nq_y, the points of local edge e generated from (v_e, v_e+1, v_e+2).
 * face_areas [F], face_normals [F,3] as in the reference.
 * zf: [3F, ldzf] complex, row = test face-edge, col = source face-edge. */
int mba_mom_zmat_faceedge_ref(int prec, const void *verts, const int32_t *faces,
                              const void *face_areas, const void *face_normals, const void *gp_x,
                              const void *gw_x, int32_t nq_x, const void *gp_y, const void *gw_y,
                              int32_t nq_y, double wavenumber, int64_t n_verts, int64_t n_faces,
                              void *zf, int64_t ldzf, mba_stream_t stream);

/* Device memory of mba_mom_zmat_faceedge_ref_noalloc per GPU (bytes). */
#define MBA_NOALLOC_SCRATCH_BYTES (4u << 20)

/* mba_mom_zmat_faceedge_ref without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and result as mba_mom_zmat_faceedge_ref; the
 * per-face records it needs (96 B per face in MBA_F32, 192 B in MBA_F64) live
 * in a fixed device array of MBA_NOALLOC_SCRATCH_BYTES per GPU, so n_faces is
 * at most 43 690 (MBA_F32) / 21 845 (MBA_F64); larger meshes return
 * MBA_ERR_SIZE before any launch.  The array is shared by every call on one
 * GPU: calls that overlap on one GPU (different streams or host threads
 * without ordering) race on it.  Every record is written before it is read,
 * so nothing needs clearing between calls.  The stream must belong to the
 * current device.  Calls no hipMalloc, hipFree or allocating library routine. */
int mba_mom_zmat_faceedge_ref_noalloc(int prec, const void *verts, const int32_t *faces,
                                      const void *face_areas, const void *face_normals, const void *gp_x,
                                      const void *gw_x, int32_t nq_x, const void *gp_y, const void *gw_y,
                                      int32_t nq_y, double wavenumber, int64_t n_verts, int64_t n_faces,
                                      void *zf, int64_t ldzf, mba_stream_t stream);

/* Face-edge RHS, reference incidence: replaces create_MOM_rhs_3d_forward
 * (mesh_MOM_3d.cpp:15-36, kernel mesh_MOM_3d.cu:244-330).  vf: [F, 3] complex. */
int mba_mom_rhs_faceedge_ref(int prec, const void *verts, const int32_t *faces,
                             const void *face_areas, const void *face_normals, const void *gp,
                             const void *gw, int32_t nq, double wavenumber, int64_t n_verts,
                             int64_t n_faces, void *vf, mba_stream_t stream);

/* RWG coefficients to face-vertex currents: replaces
 * interpolate_Icoeff_to_Js_3d_forward (mesh_MOM_3d.cpp:90-109,
 * kernel mesh_MOM_3d.cu:498-565).  icoef [B, E] complex, out [B, 3F, 3]
 * complex (overwritten, not accumulated). */
int mba_mom_interp_faceedge(int prec, const void *verts, const int32_t *faces,
                            const int32_t *face_edge_indices, const void *face_areas,
                            const void *icoef, int64_t batch, int64_t n_edges, int64_t n_verts,
                            int64_t n_faces, void *out, mba_stream_t stream);

/* Deterministic vertex accumulation (replaces the scatter_add_ of
 * mesh_BEM_solver_3d.py:186-192): vert[b, v, :] = sum over the (face, slot)
 * pairs of vertex v, in the order of vinc_idx.  vinc_ptr [V+1] int64,
 * vinc_idx [3F] int32 (= 3 f + t).  fvj [B, 3F, 3], vert [B, V, 3] complex. */
int mba_mom_vertex_sum(int prec, const void *fvj, const int64_t *vinc_ptr,
                       const int32_t *vinc_idx, int64_t batch, int64_t n_verts, int64_t n_faces,
                       void *vert, mba_stream_t stream);

/* Bytes of device workspace the ns-rule calls need for n_faces faces. */
size_t mba_mom_workspace_bytes(int prec, int64_t n_faces);

/* Face-edge EFIE blocks, NS rule: the 7-point degree-5 triangle rule with
 * analytic FP64 1/R terms on near pairs (no reference counterpart; the RWG
 * space result replaces the Zmn of mesh_BEM_solver_3d.py:128-147):
 *   Z_ij = c_i c_j [ jk <(x-v_i).(y-v_j) G> - (4j/k) <G> ],  G = e^{-jkR}/(4 pi R)
 * rule: [nq, 4] double (l0, l1, l2, w), physical weight w * 2A.
 * near_ptr [F+1] int64 / near_idx int32: CSR over source faces b of the test
 * faces a whose pair takes the singular split; for those the 1/(4 pi R) part
 * of the inner integral is computed analytically in FP64.
 * zf: [3F, ldzf] complex rows = test face-edge; cols [col0, col0 + 3*nb)
 * receive source faces [b0, b0 + nb).  The near relation must be symmetric
 * (a near b <=> b near a; mba_near_pairs_* build it so): in FP64, pairs of two
 * faces in [b0, b0 + nb) are evaluated once and stored in both orders.
 * workspace >= mba_mom_workspace_bytes. */
int mba_mom_zmat_faceedge_ns(int prec, const void *verts, const int32_t *faces, int64_t n_verts,
                             int64_t n_faces, const double *rule, int32_t nq, double wavenumber,
                             const int64_t *near_ptr, const int32_t *near_idx, int64_t b0,
                             int64_t nb, void *zf, int64_t ldzf, void *workspace,
                             size_t workspace_bytes, mba_stream_t stream);

/* Face-edge plane-wave RHS, NS rule:
 *   vf[3f+e, d] = sum_i w_i c_e (x_i - v_e).p_d e^{-jk khat_d.x_i}
 * dirs, pols: [ndir, 3] double (unit propagation and real polarisation).
 * vf: [3F, ndir] complex. */
int mba_mom_rhs_faceedge_ns(int prec, const void *verts, const int32_t *faces, int64_t n_verts,
                            int64_t n_faces, const double *rule, int32_t nq, double wavenumber,
                            const double *dirs, const double *pols, int64_t ndir, void *vf,
                            mba_stream_t stream);

/* RWG accumulation of face-edge blocks into RWG space (replaces the
 * scatter_add_ of mesh_BEM_solver_3d.py:141-147 with a fixed-order gather):
 *   z[m, n] (=|+=) sum_{s,t} zf[fe_s(m), fe_t(n) - col0],  rwg_fe [E, 2] ascending,
 * summed as (zf[fe0,fe0] + zf[fe1,fe0]) + (zf[fe0,fe1] + zf[fe1,fe1]).
 * Only source face-edges in [col0, col0 + ncols) are read: a column n whose
 * two face-edges both lie in the window is stored; one whose fe_0 lies in it
 * is stored; one whose fe_1 alone lies in it is added (its fe_0 window must
 * have been gathered before).  z: [E, ldz] complex. */
int mba_mom_gather_rwg(int prec, const void *zf, int64_t ldzf, int64_t col0, int64_t ncols,
                       const int32_t *rwg_fe, int64_t n_edges, void *z, int64_t ldz,
                       mba_stream_t stream);

/* v[m, d] = vf[fe_0(m), d] + vf[fe_1(m), d]; vf [3F, ndir], v [E, ndir] complex. */
int mba_mom_gather_rhs(int prec, const void *vf, const int32_t *rwg_fe, int64_t n_edges,
                       int64_t ndir, void *v, mba_stream_t stream);

/* Several meshes in one assembly launch (NS EFIE rule).  verts / faces
 * are the concatenation of the meshes (faces hold indices into the
 * concatenated verts; mesh m owns faces [face_off, face_off + n_faces), face_off
 * equal to the running sum); every mesh writes its own face-edge block
 * zf [3 n_faces, ldzf] for all its source faces.  n_near = near_ptr[n_faces];
 * each mesh's near relation must be symmetric (see mba_mom_zmat_faceedge_ns).
 * The descriptor array is host memory (staged by the call; it may be released
 * on return).  No reference counterpart (the reference solves one mesh per
 * call). */
typedef struct {
    int64_t n_faces;
    int64_t face_off;
    double wavenumber;
    const int64_t *near_ptr;
    const int32_t *near_idx;
    void *zf;
    int64_t ldzf;
    int64_t n_near;
} mba_ns_mesh;
size_t mba_mom_batched_workspace_bytes(int prec, int64_t total_faces, int32_t n_meshes);
int mba_mom_zmat_faceedge_ns_batched(int prec, const void *verts, const int32_t *faces, int64_t total_faces,
                                     const double *rule, int32_t nq, const mba_ns_mesh *meshes,
                                     int32_t n_meshes, void *workspace, size_t workspace_bytes,
                                     mba_stream_t stream);

/* Plane-wave right-hand sides of several meshes in one pass (NS rule,
 * as mba_mom_rhs_faceedge_ns followed by mba_mom_gather_rhs per mesh): verts /
 * faces concatenated as for mba_mom_zmat_faceedge_ns_batched (face_off the
 * running sum of n_faces); mesh m's directions and polarisations dirs / pols
 * [ndir, 3] (double), its RWG table rwg_fe [n_edges, 2] (mesh-local face-edge
 * indices 3 f + e) and its output v [n_edges, ndir] complex (prec).  Replaces
 * the per-mesh create_MOM_rhs_3d_forward calls of the reference's batch loop
 * (demo/demo_Helmholtz_MOM_3d.py:206-216 -> mesh_MOM_3d.cpp:3-36) with the
 * NS incidence; the descriptor array is host memory (staged). */
typedef struct {
    int64_t n_faces;
    int64_t face_off;
    double wavenumber;
    const double *dirs;
    const double *pols;
    int64_t ndir;
    const int32_t *rwg_fe;
    int64_t n_edges;
    void *v;
} mba_ns_rhs_mesh;
size_t mba_mom_rhs_ns_batched_workspace_bytes(int prec, int64_t total_faces, const mba_ns_rhs_mesh *meshes,
                                              int32_t n_meshes);
int mba_mom_rhs_ns_batched(int prec, const void *verts, const int32_t *faces, int64_t total_faces,
                           const double *rule, int32_t nq, const mba_ns_rhs_mesh *meshes, int32_t n_meshes,
                           void *workspace, size_t workspace_bytes, mba_stream_t stream);

/* RWG gathers of several meshes in one launch: job q takes the full
 * face-edge block zf [3F, ldzf] of its mesh (every source face, as
 * mba_mom_zmat_faceedge_ns_batched writes it) to z [n_edges, ldz], the same
 * sums in the same order as mba_mom_gather_rwg (bit-identical).  Replaces the
 * RWG scatter of mesh_BEM_solver_3d.py:128-164 for a batch; the job array is
 * host memory (staged); scratch: mba_mom_gather_rwg_batched_scratch_bytes. */
typedef struct {
    const void *zf;
    int64_t ldzf;
    const int32_t *rwg_fe;
    int64_t n_edges;
    void *z;
    int64_t ldz;
} mba_gather_job;
size_t mba_mom_gather_rwg_batched_scratch_bytes(int32_t n_jobs);
int mba_mom_gather_rwg_batched(int prec, const mba_gather_job *jobs, int32_t n_jobs, void *scratch,
                               size_t scratch_bytes, mba_stream_t stream);

/* =================================================================== dense
 * Row-major complex LU with partial (row) pivoting, in place, replacing the
 * getrf inside torch.linalg.solve (mesh_BEM_solver_3d.py:174, 346).
 * ipiv [n] int32 (0-based, LAPACK order: rows j and ipiv[j] swapped at step j).
 * info: device int32, set to 0 or 1 + (first zero pivot column).
 * workspace >= mba_lu_workspace_bytes(prec, n). */
size_t mba_lu_workspace_bytes(int prec, int64_t n);
int mba_lu_factor(int prec, void *a, int64_t n, int64_t lda, int32_t *ipiv, int32_t *info,
                  void *workspace, size_t workspace_bytes, mba_stream_t stream);

/* mba_lu_factor with per-phase device timing (HIP events on `stream`):
 * phase_ms[4] / phase_launches[4] = total milliseconds / launches of
 * 0 panel factorisation, 1 row interchanges, 2 triangular solve of the U row
 * block, 3 trailing MFMA GEMM; gemm_flops = real flops of the trailing GEMMs
 * (8 m n k per launch).  No counterpart in the reference (measurement aid). */
/* As mba_lu_factor without any host synchronisation: the cooperative-wait
 * flag (nonzero: a panel wait timed out) is copied to err_out (device int32)
 * on the stream, and each panel uses at most max_wg workgroups (0: no limit),
 * so that factorisations on several streams can run side by side.  No
 * reference counterpart (the reference factors one matrix at a time). */
int mba_lu_factor_async(int prec, void *a, int64_t n, int64_t lda, int32_t *ipiv, int32_t *info,
                        void *workspace, size_t workspace_bytes, int32_t max_wg, int32_t *err_out,
                        mba_stream_t stream);
/* Several factorisations in lock step (one matrix per entry, each with the
 * workspace of mba_lu_workspace_bytes for its order): the panels of all
 * matrices are factored together in shared launches, so the per-column
 * exchange latency is shared.  scratch: mba_lu_batched_scratch_bytes; the
 * call copies its panel schedule there and waits for that copy.  err_out:
 * device int32 [n_mat], the cooperative-wait flag of each factorisation.  No
 * reference counterpart. */
typedef struct {
    void *a;
    int64_t n;
    int64_t lda;
    int32_t *ipiv;
    int32_t *info;
    void *workspace;
    size_t workspace_bytes;
} mba_lu_matrix;
/* Many independent products C_j = alpha_j A_j B_j + beta_j C_j (row-major
 * complex, the layout of mba_gemm) in one launch: the tiles of all jobs form
 * one grid (FP32: 3M product on MFMA; FP64: MFMA).  The job table is copied to
 * scratch (mba_gemm_grouped_scratch_bytes) and the call waits for that copy.
 * Jobs must not write overlapping C.  No reference counterpart. */
typedef struct {
    const void *a;
    const void *b;
    void *c;
    int64_t m, n, k, lda, ldb, ldc;
    double alpha, beta;
} mba_gemm_job;
size_t mba_gemm_grouped_scratch_bytes(int32_t n_jobs);
int mba_gemm_grouped(int prec, const mba_gemm_job *jobs, int32_t n_jobs, void *scratch, size_t scratch_bytes,
                     mba_stream_t stream);

/* mba_lu_solve_prepare for several factorisations in two launches (the
 * permutations, then every inverted diagonal block).  The job table is copied
 * to scratch (mba_lu_solve_prepare_batched_scratch_bytes) and the call waits
 * for that copy.  No reference counterpart. */
typedef struct {
    const void *lu;
    int64_t n;
    int64_t lda;
    const int32_t *ipiv;
    void *aux;
    size_t aux_bytes;
} mba_lu_prepare_job;
size_t mba_lu_solve_prepare_batched_scratch_bytes(int32_t n_jobs);
int mba_lu_solve_prepare_batched(int prec, const mba_lu_prepare_job *jobs, int32_t n_jobs, void *scratch,
                                 size_t scratch_bytes, mba_stream_t stream);

/* Solves of several prepared factorisations (mba_lu_solve_prepare) with many
 * right-hand sides each, all in two launches (forward L, backward U; one
 * workgroup per job and 16 right-hand sides).  b [n, ldb] is overwritten with
 * the solution; work: n * nrhs complex elements per job.  The call copies the
 * job array to scratch (mba_lu_solve_batched_scratch_bytes) and waits for
 * that copy.  No reference counterpart. */
typedef struct {
    const void *lu;
    int64_t n;
    int64_t lda;
    const void *aux;
    void *b;
    int64_t ldb;
    int64_t nrhs;
    void *work;
} mba_lu_solve_job;
size_t mba_lu_solve_batched_scratch_bytes(int32_t n_jobs);
int mba_lu_solve_batched(int prec, const mba_lu_solve_job *jobs, int32_t n_jobs, void *scratch,
                         size_t scratch_bytes, mba_stream_t stream);
/* The same solves (complex64) in ONE persistent launch (all workgroups
 * resident, checked against the device's occupancy before the launch), as a
 * dataflow inside each XCD: every job goes to one XCD's queue, every 64-row
 * block and group of 64 right-hand sides of a job is a forward and a backward
 * task, taken in ticket order by that XCD's workgroups; a task subtracts the
 * products of the blocks it depends on as their flags come up, then applies
 * the diagonal block's inverse and raises its own flag.  Same jobs and
 * results as mba_lu_solve_batched; scratch:
 * mba_lu_solve_batched_persistent_scratch_bytes(jobs, n_jobs) (reads n and
 * nrhs of each job).  err_out (device int32) is written by the launch: zero,
 * or nonzero after a wait that timed out (the results are then invalid).
 * complex128 runs the two-launch path and writes zero.  Replaces the
 * per-mesh torch.linalg.solve of src/mesh_BEM/mesh_BEM_solver_3d.py:174 for
 * many right-hand sides; no single reference call. */
size_t mba_lu_solve_batched_persistent_scratch_bytes(const mba_lu_solve_job *jobs, int32_t n_jobs);
int mba_lu_solve_batched_persistent(int prec, const mba_lu_solve_job *jobs, int32_t n_jobs, void *scratch,
                                    size_t scratch_bytes, int32_t *err_out, mba_stream_t stream);
size_t mba_lu_batched_scratch_bytes(int prec, int32_t n_mat, int64_t n_max);
int mba_lu_factor_batched(int prec, const mba_lu_matrix *mats, int32_t n_mat, void *scratch, size_t scratch_bytes,
                          int32_t *err_out, mba_stream_t stream);
/* The same factorisations (complex64) in ONE persistent launch (all
 * workgroups resident, checked against the device's occupancy before the launch):
 * resident workgroups step through a table of phases separated by grid
 * barriers (32-column panels of every matrix together, left-looking super
 * panels of 256 columns, one trailing product per super panel).  Same
 * arguments and results as mba_lu_factor_batched (ipiv with LAPACK's partial
 * pivoting); scratch: mba_lu_batched_persistent_scratch_bytes.  complex128
 * runs the lock-step path.  Replaces the per-mesh loop of solves in
 * demo/demo_Helmholtz_MOM_3d.py:206-216 (each a torch.linalg.solve,
 * src/mesh_BEM/mesh_BEM_solver_3d.py:174); no single reference call. */
size_t mba_lu_batched_persistent_scratch_bytes(int prec, int32_t n_mat, int64_t n_max);
int mba_lu_factor_batched_persistent(int prec, const mba_lu_matrix *mats, int32_t n_mat, void *scratch,
                                     size_t scratch_bytes, int32_t *err_out, mba_stream_t stream);
/* Schedule of the persistent batched LU, per call: super_panels = 32-column
 * panels per super panel, 1..8 (0: 8; the trailing product's k is 32 x
 * super_panels); trailing = 0 for the FP32-MFMA 3M trailing product, 1 for the
 * FP32-accurate split-FP16 product (operand records in each matrix's LU
 * workspace).  No reference counterpart (what the batch autotuner picks). */
typedef struct {
    int32_t super_panels;
    int32_t trailing;
} mba_batched_lu_options;
/* mba_lu_factor_batched_persistent with explicit options (complex64 only). */
int mba_lu_factor_batched_persistent_opts(int prec, const mba_lu_matrix *mats, int32_t n_mat,
                                          const mba_batched_lu_options *opts, void *scratch, size_t scratch_bytes,
                                          int32_t *err_out, mba_stream_t stream);
/* mba_lu_factor_batched_persistent (complex64) with phase timing from the
 * launch's own wall clock (the call synchronises): per phase kind (0 init,
 * 1 products, 2 panels, 3 row interchanges, 4 triangular solves, 5 final)
 * kind_ms[kind] = summed work time (from the last workgroup's release by the
 * previous barrier to the last workgroup's end of work), kind_ms[6 + kind] =
 * summed grid-barrier time, kind_ms[12 + i] = the products' work time split
 * into panel-column updates (i = 0), U-row updates (1) and trailing products
 * (2), kind_ms[15] = 0; kind_phases[kind] = phase count.  No reference
 * counterpart. */
int mba_lu_factor_batched_persistent_profiled(const mba_lu_matrix *mats, int32_t n_mat,
                                              const mba_batched_lu_options *opts, void *scratch,
                                              size_t scratch_bytes, int32_t *err_out, double *kind_ms,
                                              int64_t *kind_phases, mba_stream_t stream);
/* mba_lu_factor with per-phase timing (HIP events on the stream; the call
 * synchronises): phase_ms[5] and phase_launches[5] for the panels, row
 * interchanges, triangular solves, FP32 / FP64 MFMA products (with the operand
 * splits of the split-FP16 products) and the split-FP16 trailing-update
 * products; gemm_flops[2] = algorithmic flops (8 m n k per complex product)
 * of the FP32 / FP64 products and of the split-FP16 products.  No reference
 * counterpart. */
int mba_lu_factor_profiled(int prec, void *a, int64_t n, int64_t lda, int32_t *ipiv, int32_t *info,
                           void *workspace, size_t workspace_bytes, double *phase_ms,
                           int64_t *phase_launches, double *gemm_flops, mba_stream_t stream);

/* Row permutation of a factorisation: row i of P A is row perm[i] of A
 * (the net effect of the sequential swaps in ipiv). */
int mba_lu_pivots_to_perm(const int32_t *ipiv, int64_t n, int32_t *perm, mba_stream_t stream);

/* Solve aid computed once per factorisation: the row permutation and the
 * inverses of the 64 x 64 diagonal blocks of L and U.  aux >=
 * mba_lu_solve_aux_bytes(prec, n). */
size_t mba_lu_solve_aux_bytes(int prec, int64_t n);
int mba_lu_solve_prepare(int prec, const void *lu, int64_t n, int64_t lda, const int32_t *ipiv,
                         void *aux, size_t aux_bytes, mba_stream_t stream);

/* Solve A x = b in place for nrhs right-hand sides b [n, ldb] complex, with lu
 * from mba_lu_factor and aux from mba_lu_solve_prepare (the getrs inside
 * torch.linalg.solve).  workspace >= mba_lu_solve_workspace_bytes. */
size_t mba_lu_solve_workspace_bytes(int prec, int64_t n, int64_t nrhs);
int mba_lu_solve(int prec, const void *lu, int64_t n, int64_t lda, const void *aux, void *b,
                 int64_t nrhs, int64_t ldb, void *workspace, size_t workspace_bytes,
                 mba_stream_t stream);

/* Complex GEMM C = alpha A B + beta C, row-major, A [m, k], B [k, n], C [m, n];
 * alpha, beta real.  Used by the LU; exported for tests. */
int mba_gemm(int prec, int64_t m, int64_t n, int64_t k, double alpha, const void *a, int64_t lda,
             const void *b, int64_t ldb, double beta, void *c, int64_t ldc, mba_stream_t stream);

/* Tile variants of the FP32 GEMM (0: 128x64 complex tile, 1: 128x64 ordered
 * for the 8 XCDs, 2: 128x128, 3: 128x128 XCD-ordered, 4: 128x64 with three
 * real products per complex product (3M), 5: 3M XCD-ordered; 4 and 5 switch
 * to 64x64 tiles, bit-identical, when 128-row tiles would give fewer than two
 * workgroups per CU).  mba_gemm_variant runs
 * one variant; mba_gemm_set_variant selects the one mba_gemm and the LU use
 * (process-wide, default 4; set once by a tuner, not per call).  No
 * counterpart in the reference. */
int mba_gemm_variant(int prec, int variant, int64_t m, int64_t n, int64_t k, double alpha,
                     const void *a, int64_t lda, const void *b, int64_t ldb, double beta, void *c,
                     int64_t ldc, mba_stream_t stream);
int mba_gemm_set_variant(int prec, int variant);

/* Complex64 GEMM C = alpha A B + beta C (row-major, 1 <= k <= 512) with FP32
 * accuracy on the FP16 matrix cores: rows of A and columns of B are scaled by
 * powers of two, every value is split into two FP16 pieces (hi + lo), and the
 * three piece products hi.hi + hi.lo + lo.hi accumulate in FP32 (error about
 * 2^-22 |A||B| against 2^-24 for an FP32 product).  workspace >=
 * mba_gemm_split_f16_workspace_bytes.  It is the trailing-update product of the
 * complex64 LU when mba_lu_set_update(1) is selected: a dense contraction
 * inside the getrf behind torch.linalg.solve
 * (src/mesh_BEM/mesh_BEM_solver_3d.py:174); no direct reference counterpart. */
size_t mba_gemm_split_f16_workspace_bytes(int64_t m, int64_t n, int64_t k);
int mba_gemm_split_f16(int64_t m, int64_t n, int64_t k, double alpha, const void *a, int64_t lda,
                       const void *b, int64_t ldb, double beta, void *c, int64_t ldc, void *workspace,
                       size_t workspace_bytes, mba_stream_t stream);

/* Workgroup tile of mba_gemm_split_f16 (and of the LU's split update): 0 =
 * 256 x 128 complex with 8 waves (one workgroup per CU, operand records
 * staged by direct global-to-LDS loads; default), 1 = 128 x 128 with 4 waves
 * (two workgroups per CU, register-staged); bit-identical results.
 * Process-wide tuning knob, like
 * mba_gemm_set_variant.  No counterpart in the reference. */
int mba_gemm_split_f16_set_tile(int tile);

/* Trailing-update product of the complex64 LU (mba_lu_factor and its
 * asynchronous / profiled forms): 0 = FP32 MFMA (the mba_gemm_set_variant
 * choice, default), 1 = the split-FP16 product of mba_gemm_split_f16 for
 * trailing blocks of at least 2048 x 2048.  Process-wide; set before
 * factorising, not concurrently with one.  No counterpart in the reference. */
int mba_lu_set_update(int kind);

/* Panels of the complex64 LU (mba_lu_factor and its asynchronous / profiled
 * forms): 0 = partial pivoting in LAPACK's row order (default), 1 =
 * tournament pivoting for 64-column panels of 2048 .. 128 x (compute units)
 * rows: each workgroup nominates rows by partial pivoting on its own rows, a
 * binary tree of partial-pivoting eliminations over the nominees picks the 64
 * pivot rows (about ten cross-workgroup exchanges per panel instead of two
 * per column), ipiv holds the equivalent LAPACK interchanges.  The pivot
 * order differs from LAPACK's; |L| may exceed 1 (mildly), so use it with
 * iterative refinement.  Process-wide; set before factorising.  No
 * counterpart in the reference (its getrf is the one inside
 * torch.linalg.solve, src/mesh_BEM/mesh_BEM_solver_3d.py:174). */
int mba_lu_set_panel(int kind);

/* Triangular solves of mba_lu_solve for complex64 and nrhs <= 4: 1 = one
 * persistent launch per triangle (default; workgroups take 64-row blocks in
 * start order and read solved blocks from epoch-tagged words), 0 = one launch
 * per 64-row block.  Process-wide; set before solving, not concurrently with
 * one.  Both give the getrs inside torch.linalg.solve
 * (src/mesh_BEM/mesh_BEM_solver_3d.py:174); no counterpart of the switch in
 * the reference. */
int mba_lu_set_solve(int kind);

/* FP64 residual r = v - Z x for complex128 Z [n, ldz], x/v/r [n, nrhs]. */
int mba_residual_c128(const void *z, int64_t n, int64_t ldz, const void *x, const void *v,
                      void *r, int64_t nrhs, mba_stream_t stream);

/* y = A x for complex A [n, lda] (prec), x, y [n] complex (prec). */
int mba_gemv(int prec, const void *a, int64_t n, int64_t lda, const void *x, void *y,
             mba_stream_t stream);

/* Element conversion complex128 <-> complex64 of an [m, n] matrix with leading dims. */
int mba_convert_c128_c64(const void *src, int64_t m, int64_t n, int64_t lds, void *dst,
                         int64_t ldd, mba_stream_t stream);

/* Many complex128 -> complex64 conversions (contiguous arrays of `count`
 * complex elements) in one launch; scratch:
 * mba_convert_c128_c64_batched_scratch_bytes(n_jobs).  Part of the batched
 * mixed-precision refinement; no counterpart in the reference. */
typedef struct {
    const void *src; /* complex128 */
    void *dst;       /* complex64 */
    int64_t count;
} mba_convert_job;
size_t mba_convert_c128_c64_batched_scratch_bytes(int32_t n_jobs);
int mba_convert_c128_c64_batched(const mba_convert_job *jobs, int32_t n_jobs, void *scratch,
                                 size_t scratch_bytes, mba_stream_t stream);

/* Refinement update of many systems in one pass: J += D (complex128 J
 * [n, ldj], complex64 D [n, ldd], nrhs columns each), and ratio_out[q]
 * (device double) = max over the columns c of job q of |D_c| / |J_c| after
 * the update (2-norms; +inf for a zero J column with a nonzero D column, 0
 * when both are zero), summed in a fixed order (reproducible).  Scratch:
 * mba_refine_update_batched_scratch_bytes(jobs, n_jobs) (reads nrhs of each
 * job).  The step J += LU32 \ (V - Z J) of the reference-free mixed-precision
 * refinement that replaces torch.linalg.solve
 * (src/mesh_BEM/mesh_BEM_solver_3d.py:174) for many meshes; no counterpart
 * in the reference. */
typedef struct {
    void *j; /* complex128 */
    int64_t ldj;
    const void *d; /* complex64 */
    int64_t ldd;
    int64_t n, nrhs;
} mba_refine_job;
size_t mba_refine_update_batched_scratch_bytes(const mba_refine_job *jobs, int32_t n_jobs);
int mba_refine_update_batched(const mba_refine_job *jobs, int32_t n_jobs, double *ratio_out, void *scratch,
                              size_t scratch_bytes, mba_stream_t stream);

/* ================================================================= Laplace */

/* Unpacked Laplace single (layer = 1) or double (layer = 2) layer with the
 * reference's relation-typed rule: replaces double_integration_on_panels_3d
 * (src/mesh_BEM/utils/integrations_3d.py:252-452) as called by
 * SingleLayer3d.compute_Vmat / DoubleLayer3d.compute_Kmat
 * (single_layer_3d.py:43-68, double_layer_3d.py:43-68).
 * table: rows (r1x, r2x, r1y, r2y, weight*jacobian) of the four relation types
 * concatenated, table_offsets [5] int64 (row ranges of relations 0..3).
 * out: [F, F, nphi, nphi] (row face fy, column face fx), nphi = 1 (PLANAR) or 3 (LINEAR). */
int mba_laplace_layer_ref(int prec, const void *verts, const int32_t *faces, const void *face_areas,
                          const void *face_normals, int64_t n_verts, int64_t n_faces, const void *table,
                          const int64_t *table_offsets, int32_t layer, int32_t nphi, void *out,
                          mba_stream_t stream);

/* [F, F, nphi, nphi] -> [V, V] (replaces pack_matrix_tight's scatter_add_,
 * src/utils/BEM_utils.py:314-357) with the vertex incidence CSR of
 * mba_mom_vertex_sum. */
int mba_laplace_pack(int prec, const void *unpacked, int64_t n_faces, int32_t nphi,
                     const int64_t *vinc_ptr, const int32_t *vinc_idx, int64_t n_verts, void *out,
                     mba_stream_t stream);

/* Dirichlet RHS rhs = 1/2 g .* vert_areas + K_tight^T g
 * (src/mesh_BEM/rhs_layer/rhs_layer_3d.py:122-146). */
int mba_laplace_dirichlet_rhs(int prec, const void *k_tight, int64_t n_verts, const void *g,
                              const void *vert_areas, void *rhs, mba_stream_t stream);

/* NS P1 Galerkin single and double layer (no reference counterpart:
 * the reference's layers are double_integration_on_panels_3d,
 * src/mesh_BEM/utils/integrations_3d.py:252-452, whose collapsed rule is
 * mba_laplace_layer_ref).  Unpacked layout of mba_laplace_layer_ref with
 * nphi = 3: v_unpacked / k_unpacked [F, F, 3, 3], entry (fy, fx, iy, ix) =
 * int_Tx int_Ty phi_ix(x) phi_iy(y) k(x, y), k = 1/(4 pi R) (single) or
 * (x - y).n_y / (4 pi R^3) (double, n_y the normal of face fy).  Far pairs:
 * rule [nq, 4] (barycentric l0 l1 l2, weight; nq 3 or 7) on both triangles;
 * pairs of the near CSR (a symmetric relation, mba_near_pairs_*): the rule on
 * T_x and the inner integral over T_y in closed form, in FP64.  Either output
 * may be null.  workspace: mba_laplace_p0_workspace_bytes(prec, n_faces). */
int mba_laplace_p1_assemble(int prec, const void *verts, const int32_t *faces, int64_t n_verts, int64_t n_faces,
                            const double *rule, int32_t nq, const int64_t *near_ptr, const int32_t *near_idx,
                            void *v_unpacked, void *k_unpacked, void *workspace, size_t workspace_bytes,
                            mba_stream_t stream);

/* out = alpha M g + beta out with the P1 mass matrix M_ij = int phi_i phi_j
 * (the consistent counterpart of the lumped 1/2 g vert_area of
 * RhsLayer3d.compute_rhs, src/mesh_BEM/rhs_layer/rhs_layer_3d.py:127).
 * vinc_ptr / vinc_idx: the vertex -> face-corner CSR of mba_laplace_pack. */
int mba_laplace_mass_apply(int prec, const void *verts, const int32_t *faces, const int64_t *vinc_ptr,
                           const int32_t *vinc_idx, int64_t n_verts, const void *g, double alpha, double beta,
                           void *out, mba_stream_t stream);

/* P0 Galerkin Laplace operators, NS rule (no reference counterpart):
 *   V[i, j]  = int_Ti int_Tj 1/(4 pi R)
 *   Kt[i, j] = int_Ti int_Tj n_i . grad_x 1/(4 pi R)   (adjoint double layer)
 * 7-point (or 3-point) rule on far pairs, analytic FP64 inner integrals on
 * the near pairs given by the CSR (near_ptr [F+1], near_idx) over columns j.
 * v_out / kt_out: [F, ld] (either may be null).  rule: [nq, 4] double. */
size_t mba_laplace_p0_workspace_bytes(int prec, int64_t n_faces);
int mba_laplace_p0_assemble(int prec, const void *verts, const int32_t *faces, int64_t n_verts,
                            int64_t n_faces, const double *rule, int32_t nq, const int64_t *near_ptr,
                            const int32_t *near_idx, void *v_out, void *kt_out, int64_t ld,
                            void *workspace, size_t workspace_bytes, mba_stream_t stream);

/* ================================================================== GMRES
 * Building blocks of restarted GMRES (no reference counterpart: the reference
 * solves only with torch.linalg.solve, mesh_BEM_solver_3d.py:174, 346).
 * field = MBA_COMPLEX (interleaved complex) or MBA_REAL.  Matrices and bases
 * are FP32; vprec (MBA_F32 / MBA_F64) is the precision of the vectors a call
 * reads and writes (x, b, y, w, src).  A basis is row-major [nvec, ldb]:
 * vector j at basis + j * ldb elements. */

/* y = A x, or y = b - A x when b is not null.  A: FP32 [rows, lda] row-major.
 * Products and sums in FP64. */
int mba_mv32(int field, const void *a, int64_t rows, int64_t cols, int64_t lda, int vprec,
             const void *x, const void *b, void *y, mba_stream_t stream);
/* Workspace of mba_gmres_dots for vectors of length n and up to max_vec basis vectors. */
size_t mba_gmres_workspace_bytes(int64_t n, int32_t max_vec);
/* out[j] = basis_j^H w (j < nvec), out[nvec] = (||w||^2, 0); out: double2 [nvec + 1].
 * Deterministic (fixed reduction order, FP64 across threads).  nvec <= 512. */
int mba_gmres_dots(int field, const void *basis, int64_t ldb, int32_t nvec, int vprec, const void *w,
                   int64_t n, void *out, void *workspace, size_t workspace_bytes, mba_stream_t stream);
/* w += sign * sum_j basis_j coef[j]; coef: double2 [nvec] (device).  nvec <= 512. */
int mba_gmres_update(int field, const void *basis, int64_t ldb, int32_t nvec, const void *coef,
                     double sign, int vprec, void *w, int64_t n, mba_stream_t stream);
/* dst (FP32) = src / sqrt(nrm2[0].re) (zero vector if the norm is zero); nrm2: double2 (device). */
int mba_gmres_scale(int field, int vprec, const void *src, int64_t n, const void *nrm2, void *dst,
                    mba_stream_t stream);

/* ================================================================== mesh tables
 * Near-pair CSR over source faces b (ascending test faces a): a is near b when
 * |o_a - o_b| < eta (rho_a + rho_b), o = centroid, rho = largest vertex
 * distance from o, evaluated in FP64 (the criterion of the near-singular
 * terms; the reference has no near treatment).  Two calls: count (writes
 * near_ptr [F+1]; read near_ptr[F] for the size of near_idx), then fill, with
 * the same workspace. */
size_t mba_near_workspace_bytes(int64_t n_faces);
int mba_near_pairs_count(int prec, const void *verts, const int32_t *faces, int64_t n_verts,
                         int64_t n_faces, double eta, int64_t *near_ptr, void *workspace,
                         size_t workspace_bytes, mba_stream_t stream);
int mba_near_pairs_fill(int64_t n_faces, double eta, const int64_t *near_ptr, int32_t *near_idx,
                        const void *workspace, size_t workspace_bytes, mba_stream_t stream);
/* a[i, i] += alpha * d[i] for a real [n, lda] matrix (prec selects float / double). */
int mba_add_diag(int prec, void *a, int64_t n, int64_t lda, double alpha, const void *d,
                 mba_stream_t stream);

/* ================================================================== mesh-graph network
 * Kernels of the learned warm start (no reference counterpart).  GEMM
 * operands are rounded to BF16 when staged; products accumulate in FP32
 * (v_mfma_f32_32x32x8bf16); all HBM data is FP32.
 *
 * Virtual row-major operand: the concatenation of nseg (<= 3) column segments,
 * segment q holding w[q] columns of p[q] (row stride ld[q]) at row idx[q][r]
 * (or r when idx[q] is null), followed by one column of ones when ones != 0. */
typedef struct {
    const float *p[3];
    const int32_t *idx[3];
    int64_t ld[3];
    int32_t w[3];
    int32_t nseg;
    int32_t ones;
} mba_mgn_vmat;

/* Y[M, N] = epilogue(Aop Bop): Aop[m, k] = a_trans ? A(k, m) : A(m, k);
 * Bop[k, n] = b_trans ? B[n ldb + k] : B[k ldb + n].  Epilogue, in order:
 * v *= silu'(Zp[m, n]) (Zp not null); Z[m, n] = v (Z not null); v = silu(v)
 * (act = 1); v += R[m, n] (R not null); Y = v or Y += v (accumulate).  Z, R,
 * Zp share ldy.  split_k > 1 splits K over blocks into partial [split_k, M, N]
 * (caller-allocated) reduced in a fixed order; epilogue options other than
 * accumulate are then not allowed. */
int mba_mgn_gemm(const mba_mgn_vmat *a, int a_trans, const float *b, int64_t ldb, int b_trans, int64_t m,
                 int64_t n, int64_t k, float *y, int64_t ldy, float *z, const float *r, const float *zp,
                 int act, int accumulate, int32_t split_k, float *partial, mba_stream_t stream);
/* out[i, c] (=/+=) sum over q in [ptr[i], ptr[i+1]) of in[idx[q] ldin + coff + c], c < width. */
int mba_mgn_segsum(const int64_t *ptr, const int32_t *idx, int64_t rows, const float *in, int64_t ldin,
                   int32_t coff, int32_t width, float *out, int64_t ldout, int accumulate,
                   mba_stream_t stream);
/* Relative L2 loss per graph (rows [graph_off[g], graph_off[g+1]) of [*, 2] arrays)
 * and, when grad is not null, the gradient of the mean over graphs. */
int mba_mgn_loss(const float *pred, const float *target, const int64_t *graph_off, int32_t n_graphs,
                 float *loss, float *grad, mba_stream_t stream);
/* One Adam step (bias-corrected, step >= 1). */
int mba_mgn_adam(float *param, const float *grad, float *m, float *v, int64_t n, float lr, float beta1,
                 float beta2, float eps, int64_t step, mba_stream_t stream);

/* ---- field radiation (post-processing of a MoM solution) ----------------
 * Scattered magnetic field at points p (row-major [n_points, 3]) of vertex
 * currents J_v (complex [n_verts, 3]) with vertex areas A_v:
 *   out[p] = sum_v A_v J_v x grad_y G(p, v),
 *   grad_y G = e^{-jkR} (1 + jkR) (p - v) / (4 pi R^3),
 * pairs with R < 1e-6 left out.  Replaces the dense torch expression of
 * demo/demo_Helmholtz_MOM_3d.py:246-257 (gradG_y, src/utils/functions.py:844-868).
 * out: complex [n_points, 3].  The workspace (mba_radiate_h_workspace_bytes,
 * 0 for small vertex counts) holds per-slice partial sums, added in a fixed
 * order. */
size_t mba_radiate_h_workspace_bytes(int prec, int64_t n_points, int64_t n_verts);
int mba_radiate_h(int prec, const void *points, int64_t n_points, const void *verts, const void *vert_js,
                  const void *vert_areas, int64_t n_verts, double wavenumber, void *out, void *workspace,
                  size_t workspace_bytes, mba_stream_t stream);

/* ---- ferrofluid step: surface Helmholtz-decomposition sums --------------
 * For velocity fields b < batch (vert_vel [batch, n_verts, 3]) and targets x
 * ([n_targets, 3]):
 *   out[b, x] = sum_f sum_q W_fq [ (n_f . u_fq) g + (n_f x u_fq) x g ],
 *   g = (y_fq - x) / (4 pi |y_fq - x|^3), W_fq = 2 A_f w_q r1_q,
 * y_fq, u_fq the linear interpolation at (r1_q, r2_q) on the unit panel of
 * face f; rule = [nq, 3] rows (r1, r2, w).  Replaces the u1 + u2 tensor sums of
 * MeshHelmholtzDecomposition.solve (src/mesh_ferrofluid/mesh_Hemholtz_decomposition.py:80-126).
 * out: [batch, n_targets, 3].  Workspace: mba_helmholtz_decomp_workspace_bytes
 * (per-slice partial sums, added in a fixed order; 0 when not sliced). */
size_t mba_helmholtz_decomp_workspace_bytes(int prec, int64_t n_targets, int64_t n_faces, int32_t nq,
                                            int64_t batch);
int mba_helmholtz_decomp(int prec, const void *targets, int64_t n_targets, const void *verts,
                         const int32_t *faces, int64_t n_verts, int64_t n_faces, const void *vert_vel,
                         int64_t batch, const void *rule, int32_t nq, void *out, void *workspace,
                         size_t workspace_bytes, mba_stream_t stream);

/* ---- 2-D BEM on closed curves of line segments --------------------------
 * verts [V, 2], faces [F, 2] int32 (segments), prec as everywhere.
 * mba_bem2d_layer: the unpacked Galerkin layer of the reference's 2-D solver,
 * double_integration_on_panels_2d (src/mesh_BEM/utils/integrations_2d.py:56-232;
 * called by SingleLayer2d.compute_Vmat / DoubleLayer2d.compute_Kmat,
 * single_layer_2d.py:43-68, double_layer_2d.py:43-68):
 *   out[fy, fx, iy, ix] = A_y A_x sum_{a,b} w_a w_b J phi_iy phi_ix f(x, y),
 * x on source segment fx, y on test segment fy, Gauss points gp / gw [nq]
 * (the reference's float32-rounded table), the SAME_EDGE map of
 * get_rands_and_jacobian_2d (integrations_2d.py:18-53); layer 1: f = G,
 * layer 2: f = grad_y G . n_fy; bem_type 0 LAPLACE (real output), 1 / 2
 * HELMHOLTZ_BEM / HELMHOLTZ_MOM (complex output, G_2d / grad_G_y_2d with the
 * reference's large-argument Hankel form, BEM_utils.py:9-138).  order_type 0
 * PLANAR (nphi = 1) or 1 LINEAR (nphi = 2).  out [F, F, nphi, nphi].
 * mba_bem2d_pack: pack_matrix_tight with dim = 2 (BEM_utils.py:314-357), a
 * fixed-order gather over the vertex -> slot (2 f + t) CSR; out [V, V] real or
 * complex (is_complex).
 * mba_bem2d_dirichlet_rhs: rhs[v] = c_mass g[v] va[v] + c_k sum_vy Kt[vy, v] g[vy]
 * (RhsLayer2d.compute_rhs, rhs_layer_2d.py:66-130: c_mass = 1/2, c_k = 1, or
 * c_mass = 1, c_k = 0 for HELMHOLTZ_MOM); g, Kt, rhs real or complex, va real. */
int mba_bem2d_layer(int prec, int bem_type, int layer, int order_type, const void *verts, const int32_t *faces,
                    int64_t n_verts, int64_t n_faces, const void *gp, const void *gw, int32_t nq, double wavenumber,
                    void *out, mba_stream_t stream);
int mba_bem2d_pack(int prec, int is_complex, const void *unpacked, int64_t n_faces, int32_t nphi,
                   const int64_t *vinc_ptr, const int32_t *vinc_idx, int64_t n_verts, void *out, mba_stream_t stream);
int mba_bem2d_dirichlet_rhs(int prec, int is_complex, const void *k_tight, int64_t n_verts, const void *g,
                            const void *vert_areas, double c_mass, double c_k, void *rhs, mba_stream_t stream);

/* ---- ferrofluid pressure step around the Laplace solve -------------------
 * MeshPressure.solve (src/mesh_ferrofluid/mesh_pressure.py:54-163) without
 * its BEM solve (mba_laplace_* / MeshBEMSolver3d.solve).  Vertex -> face-slot
 * CSR vinc_ptr [V+1] int64 / vinc_idx [3F] int32 (slot 3 f + t, ascending).
 * mba_pressure_bc: vertex normals normalise(sum_f A_f n_f) [V, 3]; signed
 * mean curvature kappa [V] of pytorch3d's cot_laplacian (mesh_pressure.py:111-124:
 * lv = 0.5 sum_i L_ij (v_i - v_j) / sum_f A_f, kappa = 0.5 |lv| sign(lv . n));
 * p_dirichlet [B, V] = rho g v_y - P_mag + sigma kappa (0 at NEUMANN
 * vertices; p_mag [B, V] or null); dpdn_neumann [B, V] = (rho / dt) vel . n
 * (0 at DIRICHLET vertices); flags [B, V] uint8 (0 DIRICHLET, 1 NEUMANN),
 * vel [B, V, 3].  mba_pressure_update: new_vel [B, V, 3] =
 * vel - (dt / rho) (dpdn n + tangential gradient of p)
 * (compute_mesh_vert_tangent_gradient, mesh_utils.py:501-615). */
int mba_pressure_bc(int prec, const void *verts, const int32_t *faces, int64_t n_verts, int64_t n_faces,
                    const int64_t *vinc_ptr, const int32_t *vinc_idx, const void *vel, const uint8_t *flags,
                    const void *p_mag, int64_t batch, double density, double gravity_strength,
                    double surface_tension, double dt, void *vert_normals, void *kappa, void *p_dirichlet,
                    void *dpdn_neumann, mba_stream_t stream);
int mba_pressure_update(int prec, const void *verts, const int32_t *faces, int64_t n_verts, int64_t n_faces,
                        const int64_t *vinc_ptr, const int32_t *vinc_idx, const void *vert_normals, const void *p,
                        const void *dpdn, const void *vel, int64_t batch, double density, double dt, void *new_vel,
                        mba_stream_t stream);

/* ---- mesh topology on the device -----------------------------------------
 * Unique edges of a triangle mesh (replaces get_edge_unique,
 * src/utils/mesh_utils.py:921-977; local edge e of face f joins
 * faces[f][(e+1)%3] and faces[f][(e+2)%3], key max * V + min).  The
 * reference's torch.unique(sorted=False) order is backend-dependent; here
 * edges ascend by key.  Outputs (device, capacity 3 n_faces edges):
 * edges [E, 2] (max, min), face_edge_indices [3 n_faces] (edge of slot 3f+e),
 * edge_adjacent_faces [E, 2] (first, second face by ascending slot; -1 if the
 * edge has one face), *n_edges (device int64) = E.  workspace >=
 * mba_edge_unique_workspace_bytes (0 when the sizes are out of range). */
size_t mba_edge_unique_workspace_bytes(int64_t n_faces, int64_t n_verts);
int mba_edge_unique(const int32_t *faces, int64_t n_faces, int64_t n_verts, int32_t *edges,
                    int32_t *face_edge_indices, int32_t *edge_adjacent_faces, int64_t *n_edges, void *workspace,
                    size_t workspace_bytes, mba_stream_t stream);

/* ---- mesh files (host only, no device pointers) --------------------------
 * mba_mesh_read replaces the readers behind LoadSingleMesh
 * (src/utils/io.py:60-139: pywavefront for .obj, open3d for .ply, json, trimesh
 * for .dae) and reads the BEAST .in layout of demo/generate_in_file.py:12-40.
 * The format follows the extension (.obj .ply .json .dae .in); polygons are
 * split into fans (p0, p_i, p_i+1); a .dae file gives one part per geometry
 * (trimesh Scene), every other format one part.  The parts array and its
 * buffers are allocated by the library: release them with mba_mesh_free.
 * verts [n_verts, 3] double, faces [n_faces, 3] int32 (0-based).
 * mba_mesh_write_beast_in replaces save_beast_in (generate_in_file.py:12-40):
 * "1", "NV NF", vertex lines (17 significant digits), 1-based face lines (the
 * indices are shifted when the smallest one is 0). */
typedef struct {
    int64_t n_verts;
    int64_t n_faces;
    double *verts;
    int32_t *faces;
} mba_mesh_part;
int mba_mesh_read(const char *path, mba_mesh_part **parts, int32_t *n_parts);
void mba_mesh_free(mba_mesh_part *parts, int32_t n_parts);
int mba_mesh_write_beast_in(const char *path, const double *verts, int64_t n_verts, const int32_t *faces,
                            int64_t n_faces);

/* Remeshing (remesh_and_transfer_velocity, src/utils/mesh_utils.py:728-815).
 * mba_loop_subdivide: one pass of Loop subdivision on the host (the
 * reference calls open3d's subdivide_loop, mesh_utils.py:786-791): verts
 * [n_verts, 3] double, faces [n_faces, 3] int32 -> new_verts [n_new_verts, 3]
 * (the moved old vertices, then one per unique edge in ascending (min, max)
 * order), new_faces [4 n_faces, 3].  Call with new_verts == NULL to get
 * n_new_verts only.  MBA_ERR_ARG for an index out of range or a face with a
 * repeated vertex.
 * mba_nearest_transfer: the GPU step (mesh_utils.py:800-815): for every
 * point, the nearest vertex (squared distance from coordinate differences;
 * ties to the smallest index, as argmin of torch.cdist) into nearest [n_points]
 * and new_vel[b][i][:] = vel[b][nearest[i]][:] for vel [batch, n_verts, 3]
 * (prec reals; batch may be 0).  workspace >=
 * mba_nearest_transfer_workspace_bytes. */
int mba_loop_subdivide(const double *verts, int64_t n_verts, const int32_t *faces, int64_t n_faces,
                       double *new_verts, int32_t *new_faces, int64_t *n_new_verts);
/* Quadric edge-collapse decimation on the host (the reference's open3d
 * simplify_quadric_decimation, mesh_utils.py:793-794; Garland-Heckbert quadrics,
 * link-condition and normal-flip checks, boundary vertices kept) down to at
 * most target_faces faces (fewer collapses when none is valid).  out_verts
 * [n_verts, 3] / out_faces [n_faces, 3] are caller buffers of the input
 * sizes; the counts used go to n_out_verts / n_out_faces (survivors in their
 * original order, renumbered). */
int mba_decimate_quadric(const double *verts, int64_t n_verts, const int32_t *faces, int64_t n_faces,
                         int64_t target_faces, double *out_verts, int32_t *out_faces, int64_t *n_out_verts,
                         int64_t *n_out_faces);
size_t mba_nearest_transfer_workspace_bytes(int prec, int64_t n_points, int64_t n_verts);
int mba_nearest_transfer(int prec, const void *points, int64_t n_points, const void *verts, int64_t n_verts,
                         const void *vel, int32_t batch, int32_t *nearest, void *new_vel, void *workspace,
                         size_t workspace_bytes, mba_stream_t stream);

/* Vector-ALU issue rates of the current device, measured (the call
 * synchronises): out[4] = operations per second of FP32 FMA (not packed),
 * v_sin_f32, v_cos_f32 and v_rsq_f32 -- the rates the FP32 assembly kernels
 * are measured against.  blocks_per_cu: workgroups of 256 threads per CU (0:
 * 8).  No reference counterpart (a measurement of the device). */
int mba_measure_valu_rates(double *out, int blocks_per_cu, mba_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* MESH_BEM_AMD_H */
