/*
 * feasm.h -- C ABI of the MI355X finite-element assembly library (libfeasm.so).
 *
 * The library replaces the FE assembly hot path of FinEtools.jl: element
 * matrices integrated at quadrature points (IntegDomain/IntegRule Jacobians,
 * FEMMBase bilinear forms for diffusion, linear elasticity and mass) and
 * scattered through SysmatAssemblerSparse into a CSC matrix.
 *
 * Conventions (those of the reference, so that a Julia `ccall` passes its
 * arrays without copies):
 *   - node numbers, DOF numbers and CSC indices are 1-based int64 at the ABI;
 *   - node coordinates are passed as Julia stores FENodeSet.xyz: an
 *     nnodes x 3 column-major Float64 matrix (xyz[n + nnodes*k]);
 *   - connectivity is nelem x npe, passed row-major (conn[e*npe + a]), which
 *     is what `reduce(vcat, fes.conn)`-style packing of the NTuple vector gives;
 *   - dofnums is NodalField.dofnums: nnodes x ndn column-major Int64;
 *   - element matrices are column-major (Julia Matrix{Float64}).
 * Every entry point returns FEASM_OK (0) or a negative status; the message
 * text of the last failure on a context is returned by feasm_last_error().
 * Messages for the error cases the reference raises carry the reference's
 * own text ("Row degree of freedom < 1", "Wrong size of matrix", ...).
 *
 * Threading: a context owns one HIP stream and is not thread-safe; use one
 * context per host thread.  Device memory is owned by the library handles and
 * released by the matching *_destroy call.
 */
#ifndef FEASM_H
#define FEASM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define FEASM_OK 0
#define FEASM_EINVAL -1     /* bad argument                                   */
#define FEASM_EDOF -2       /* DOF number out of 1..n (assemble!)             */
#define FEASM_ESIZE -3      /* element matrix size mismatch (assemble!)       */
#define FEASM_EHIP -4       /* HIP runtime failure                            */
#define FEASM_ENOMEM -5     /* device memory exhausted / admission refused    */
#define FEASM_ESTATE -6     /* call out of protocol order                     */
#define FEASM_EUNSUPPORTED -7
#define FEASM_ECANCELLED -8 /* serving: request cancelled                     */

/* ---- enumerations ------------------------------------------------------- */
/* element types: FESetH8 (src/FESetModule.jl:953), FESetT10 (:1393) */
#define FEASM_H8 0
#define FEASM_T10 1
/* bilinear forms (src/FEMMBaseModule.jl):
 *   DIFFUSION_ISO  bilform_diffusion, scalar DataCache   (:1462-1474, :1508-1535)
 *   DIFFUSION_MAT  bilform_diffusion, 3x3 DataCache      (:1476-1506)
 *   LIN_ELASTIC    bilform_lin_elastic, DeforModelRed3D  (:1774-1813)
 *   DOT            bilform_dot with m = 3                (:1335-1366)        */
#define FEASM_FORM_DIFFUSION_ISO 0
#define FEASM_FORM_DIFFUSION_MAT 1
#define FEASM_FORM_LIN_ELASTIC 2
#define FEASM_FORM_DOT 3
/* integration rules: GaussRule(3, order) (src/IntegRuleModule.jl:206-398),
 * TetRule(npts) (src/IntegRuleModule.jl:483-516) */
#define FEASM_RULE_GAUSS 0
#define FEASM_RULE_TET 1
/* arithmetic: FP64 everywhere, or FP32 element arithmetic with FP64 Jacobian
 * determinant/inverse and FP64 global summation */
#define FEASM_FP64 0
#define FEASM_MIXED 1

typedef struct feasm_ctx feasm_ctx;
typedef struct feasm_mesh feasm_mesh;
typedef struct feasm_matrix feasm_matrix;
typedef struct feasm_coo feasm_coo;
typedef struct feasm_batch feasm_batch;
typedef struct feasm_server feasm_server;

/* ---- context ------------------------------------------------------------ */
/* No reference counterpart (the reference is host-only Julia): owns the HIP
 * device selection, one stream and the last-error text. */
int feasm_ctx_create(int device, feasm_ctx **out);
int feasm_ctx_destroy(feasm_ctx *ctx);
const char *feasm_last_error(const feasm_ctx *ctx);
/* hipStream_t of the context, as an opaque pointer (for event timing). */
void *feasm_ctx_stream(feasm_ctx *ctx);
int feasm_ctx_synchronize(feasm_ctx *ctx);
/* Free and total device memory in bytes (admission control). */
int feasm_mem_info(feasm_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes);
/* Device memory pool (no reference counterpart: Julia's GC has no
 * equivalent): buffers the library releases are kept for reuse by later
 * calls and contexts on the same device -- a hipFree would wait for the whole
 * device -- until an allocation fails or this call frees them (after a
 * device synchronisation).  *freed_bytes (may be null): the bytes freed. */
int feasm_pool_trim(int device, int64_t *freed_bytes);
/* Per-kernel timing with HIP events on the context stream.  slots > 0 turns
 * it on with a ring of `slots` numeric assemblies (each records an event
 * before the element kernel, between the kernels and after the gather);
 * 0 turns it off. */
int feasm_ctx_set_timing(feasm_ctx *ctx, int slots);
/* Durations (ms) of the recorded numeric assemblies, oldest first:
 * out[3i] element kernel, out[3i+1] gather kernel, out[3i+2] their sum.
 * Waits for the recorded events. */
int feasm_timings(feasm_ctx *ctx, double *out, int max_calls, int *ncalls);
int feasm_timings_reset(feasm_ctx *ctx);

/* Measurement helper (no reference counterpart): `repeats` launches of a
 * streaming copy of nbytes; *ms_per_copy = average duration.  Used to scale
 * hardware traffic counters to bytes. */
int feasm_copy_calibration(feasm_ctx *ctx, int64_t nbytes, int repeats, double *ms_per_copy);

/* Measurement helper (no reference counterpart): throughput of the
 * per-element node-pair Gram M_ab = sum_q u_a(q) u_b(q)^T of an 8-node,
 * 8-point element (the product inside add_btdb_ut_only!,
 * src/MatrixUtilityModule.jl:189-216) from element records in LDS, on the
 * FP64 vector ALUs (variant 0) or the FP64 matrix cores
 * (v_mfma_f64_16x16x4f64, variant 1), nelem synthetic elements (a multiple of
 * 16) each computed reps times.  *ms the kernel time, *checksum a weighted
 * sum of every Gram entry; grams_host (nkeep x 36 x 9, optional) the upper
 * node-pair blocks of the first nkeep elements. */
int feasm_probe_gram(feasm_ctx *ctx, int variant, int64_t nelem, int reps, double *ms, double *checksum,
                     double *grams_host, int64_t nkeep);

/* ---- mesh + DOF numbering (device resident) ----------------------------- */
/* Replaces what FENodeSet/FESetH8/FESetT10 hold (src/FENodeSetModule.jl:26-32,
 * src/FESetModule.jl:57-93) and the geometry NodalField (src/NodalFieldModule.jl:
 * 34-44): copies conn (1-based) and xyz to the device. */
int feasm_mesh_create(feasm_ctx *ctx, int elem_type, int64_t nelem, int npe, const int64_t *conn,
                      int64_t nnodes, const double *xyz_colmajor, feasm_mesh **out);
/* Same, from device memory already resident in HBM: conn0 int32 0-based
 * row-major (nelem x npe), xyz row-major (nnodes x 3).  The mesh copies them;
 * an entry of conn0 outside [0, nnodes) is FEASM_EINVAL, as on the host path. */
int feasm_mesh_create_device(feasm_ctx *ctx, int elem_type, int64_t nelem, int npe, const int32_t *d_conn0,
                             int64_t nnodes, const double *d_xyz, feasm_mesh **out);
int feasm_mesh_destroy(feasm_mesh *mesh);
/* Update node coordinates in place (row-major nnodes x 3, host memory). */
int feasm_mesh_set_xyz(feasm_mesh *mesh, const double *xyz_rowmajor);
/* numberdofs!/gatherdofnums! (src/FieldModule.jl:304-377): dofnums nnodes x ndn
 * column-major, 1-based; nalldofs (src/FieldModule.jl:102).  When dofnums is
 * the node-major numbering (node-1)*ndn + comp, assembly uses the device
 * resident CSC pattern; otherwise it goes through the generic COO path. */
int feasm_dofs_set(feasm_mesh *mesh, int ndn, const int64_t *dofnums_colmajor, int64_t nalldofs);
/* Node-major numbering without passing the table (numberdofs! with no EBCs). */
int feasm_dofs_set_nodemajor(feasm_mesh *mesh, int ndn);

/* ---- bilinear forms (GPU element loop + assembly) ----------------------- */
/* One call = the whole bilform_* loop followed by makematrix! on a
 * SysmatAssemblerSparse(0.0): bilform_diffusion / bilform_lin_elastic /
 * bilform_dot (src/FEMMBaseModule.jl:1335-1366, 1462-1545, 1774-1813).
 * coeff: DIFFUSION_ISO 1 value (k); DIFFUSION_MAT 9 (3x3 row-major);
 * LIN_ELASTIC 36 (6x6 row-major D); DOT ndn*ndn (row-major c).
 * rule/order: FEASM_RULE_GAUSS with order 1..4 (H8), FEASM_RULE_TET with 1, 4
 * or 5 points (T10).  The first call for a (mesh, ndn) builds the CSC pattern
 * on the device (symbolic phase); later calls reuse it (numeric phase only).
 * *out receives a new matrix handle. */
int feasm_bilform(feasm_mesh *mesh, int form, int rule, int order, const double *coeff, int ncoeff, int precision,
                  feasm_matrix **out);
/* Re-assembly into an existing matrix built for the same mesh, DOFs and form
 * class (values overwritten; pattern reused; no device allocation when the
 * staging buffers already exist). */
int feasm_bilform_into(feasm_mesh *mesh, int form, int rule, int order, const double *coeff, int ncoeff,
                       int precision, feasm_matrix *mat);
/* Generic path used for non-node-major DOF numberings, and selectable for
 * testing: element kernels write COO triples in assemble!'s order, then the
 * device sparse() (radix sort + ordered duplicate summation). */
int feasm_bilform_coo(feasm_mesh *mesh, int form, int rule, int order, const double *coeff, int ncoeff,
                      int precision, feasm_matrix **out);
/* Numeric path of the resident-pattern assembly: 0 automatic (the fused
 * tile kernel where the element type and rule have one: H8 with GaussRule(3,2),
 * T10 with TetRule(4)), 1 staged element blocks + ordered gather (every
 * element/rule), 2 fused (error where unavailable).  Both give the same
 * pattern; values agree to rounding. */
int feasm_mesh_set_numeric_path(feasm_mesh *mesh, int path);
/* Tile parameters of the fused kernel for this mesh (0 = default):
 * owned nodes per tile (<= 255), LDS per workgroup in KiB, threads (256/512/1024). */
int feasm_mesh_set_tile_params(feasm_mesh *mesh, int max_nodes, int budget_kib, int threads);
/* Tiles built for the fused kernel so far (summed over forms/precisions). */
int feasm_mesh_tile_info(feasm_mesh *mesh, int64_t *ntiles, int64_t *nhalo);
/* Tile builders, for testing (no reference counterpart): builds the fused
 * kernel's tiles for this mesh, form and precision (with the mesh's tile
 * parameters) with the device builder the assembly uses and with the host
 * builder, and compares headers, blobs and halo-node lists byte for byte.
 * *ntiles: tiles of the device build; *first_diff: the first tile that
 * differs, -1 when none; *device_ms: the device build's wall time.  Needs
 * the mesh's pattern (feasm_pattern_build). */
int feasm_tiles_compare(feasm_mesh *mesh, int form, int precision, int64_t *ntiles, int64_t *first_diff,
                        double *device_ms);
/* Build (or rebuild) the device CSC pattern for the mesh's DOFs now. */
int feasm_pattern_build(feasm_mesh *mesh);
/* Pattern statistics: nnz of the DOF-level CSC and number of node blocks. */
int feasm_pattern_info(feasm_mesh *mesh, int64_t *nnz, int64_t *nblocks);

/* ---- load vectors (linform_dot + SysvecAssembler) ----------------------- */
/* linform_dot(self, SysvecAssembler, geom, P, f, m = 3) with a constant force
 * (src/FEMMBaseModule.jl:1207-1241; distribloads :1262-1297 forwards to it)
 * followed by makevector! (src/AssemblyModule.jl:853-917): per element and
 * quadrature point elvec[(k, mx)] += N_k (det J w) force[mx], then
 * F[dofnums[i]] += elvec[i] element by element on a zeroed vector.  force has
 * one value per DOF of a node (nforce == ndn).  The DOF numbers set on the
 * mesh must be distinct.  F_device: nalldofs doubles on the device, entry
 * d-1 for DOF d, written by kernels enqueued on the context's stream
 * (feasm_ctx_stream): synchronise with it before reading F_device elsewhere. */
int feasm_linform_dot_device(feasm_mesh *mesh, int rule, int order, const double *force, int nforce,
                             double *F_device);
/* The same into a host array of nalldofs doubles (makevector!'s copy). */
int feasm_linform_dot(feasm_mesh *mesh, int rule, int order, const double *force, int nforce, double *F_host);

/* ---- lumped mass (SysmatAssemblerSparseHRZLumpingSymm) ------------------ */
/* bilform_dot (src/FEMMBaseModule.jl:1335-1366) with an ndn x ndn coefficient
 * (row-major, ncoeff == ndn*ndn) into a SysmatAssemblerSparseHRZLumpingSymm
 * and makematrix! (src/AssemblyModule.jl:943-1141): each element mass
 * matrix's diagonal scaled by (sum of its entries) / (sum of its diagonal),
 * summed per DOF in element order; the diagonal CSC, one entry per DOF that
 * some element touches.  The DOF numbers set on the mesh must be distinct. */
int feasm_bilform_dot_lumped(feasm_mesh *mesh, int rule, int order, const double *coeff, int ncoeff,
                             feasm_matrix **out);

/* ---- matrices ----------------------------------------------------------- */
/* makematrix! result (src/AssemblyModule.jl:304-329): SparseMatrixCSC{Float64,Int}. */
int feasm_matrix_dims(const feasm_matrix *mat, int64_t *m, int64_t *n, int64_t *nnz);
/* Copy to caller arrays (colptr n+1, rowval nnz, nzval nnz); base 1 gives the
 * Julia SparseMatrixCSC convention, base 0 the scipy one. */
int feasm_matrix_copy_csc(feasm_matrix *mat, int base, int64_t *colptr, int64_t *rowval, double *nzval);
/* Device pointers: colptr int64 (0-based, n+1), rowval int32 (0-based),
 * nzval double.  Valid until the matrix is destroyed. */
int feasm_matrix_device_csc(feasm_matrix *mat, const int64_t **colptr, const int32_t **rowval,
                            double **nzval);
int feasm_matrix_destroy(feasm_matrix *mat);
/* No reference counterpart (verification helper): a 64-bit digest of the
 * matrix -- column pointers, row indices and the bits of every value, each
 * hashed with its position and summed modulo 2^64 on the device -- so that
 * two large results can be compared bitwise without copying them out. */
int feasm_matrix_checksum(feasm_matrix *mat, uint64_t *digest);
/* A new matrix holding mat's entries except the stored exact zeros
 * (SparseArrays.dropzeros).  bilform_* into a SysmatAssemblerSparseSymm
 * (src/AssemblyModule.jl:342-583) gives the full symmetric matrix with this
 * pattern: its makematrix! forms S + transpose(S) of the lower-triangle
 * triples, and sparse + stores no zero result. */
int feasm_matrix_dropzeros(feasm_matrix *mat, feasm_matrix **out);
/* The free-free block: matrix_blocked_ff(A, nfreedofs)
 * (src/MatrixUtilityModule.jl:675-690; SysmatAssemblerFFBlock's makematrix!,
 * src/AssemblyModule.jl:1149-1231) of the same matrix assembled with
 * numberdofs!'s free-then-fixed numbering (src/FieldModule.jl:343-377).  That
 * numbering keeps the free DOFs in their relative order, so the block is the
 * principal submatrix of mat on the DOFs with is_free[d] != 0 (one flag per
 * column of mat, 0-based), renumbered consecutively.  A new matrix. */
int feasm_matrix_free_block(feasm_matrix *mat, const uint8_t *is_free, feasm_matrix **out);
/* matrix_blocked_ff(A, row_nfreedofs, col_nfreedofs)
 * (src/MatrixUtilityModule.jl:675-690), as SysmatAssemblerFFBlock's
 * makematrix! applies it to its wrapped assembler's matrix (:1227-1229):
 * A[1:row_nfreedofs, 1:col_nfreedofs], stored entries kept; spzeros when
 * either size is 0; "The ff block has too many rows/columns" beyond A.
 * A new matrix. */
int feasm_matrix_ff_block(feasm_matrix *mat, int64_t row_nfreedofs, int64_t col_nfreedofs, feasm_matrix **out);

/* ---- generic assembler protocol ---------------------------------------- */
/* assembler kinds of feasm_coo_create_kind (src/AssemblyModule.jl) */
#define FEASM_ASM_SPARSE 0 /* SysmatAssemblerSparse               :88-329   */
#define FEASM_ASM_SYMM 1   /* SysmatAssemblerSparseSymm           :342-583  */
#define FEASM_ASM_DIAG 2   /* SysmatAssemblerSparseDiag           :599-794  */
#define FEASM_ASM_HRZ 3    /* SysmatAssemblerSparseHRZLumpingSymm :943-1141 */
/* SysmatAssemblerSparse(0.0, nomatrixresult) (src/AssemblyModule.jl:159-165).
 * ctx may be NULL for host-side buffering only (makematrix! then fails). */
int feasm_coo_create(feasm_ctx *ctx, int nomatrixresult, feasm_coo **out);
/* The constructor of the assembler of the given kind (z = 0.0):
 * SysmatAssemblerSparseSymm(0.0, nomatrixresult) (:393-405), ...Diag (:631-643),
 * ...HRZLumpingSymm (:975-987).  The protocol calls below follow each kind's
 * own methods: Symm keeps the lower triangle of each square element matrix
 * in local numbering (assemble!, :500-549) and its makematrix! forms
 * S + transpose(S) with the diagonal halved (:551-583; exact zero sums are
 * not stored, as SparseArrays' `+` stores none); Diag keeps the diagonal
 * entries (:734-760); HRZLumpingSymm keeps the diagonal scaled by the
 * element's total over its diagonal total (:1085-1114). */
int feasm_coo_create_kind(feasm_ctx *ctx, int kind, int nomatrixresult, feasm_coo **out);
/* startassembly! (src/AssemblyModule.jl:209-238, :424-454, :671-699,
 * :1031-1061): on a fresh assembler the buffers are sized for
 * expectedntriples (:54-59, :357-362, :617-622, :961-966) and the matrix
 * dimensions recorded (the Symm, Diag and HRZ kinds require them equal:
 * "Row and column info do not agree"); on a used one both are left as they
 * are and only the buffer pointer is reset.  assemble! grows the buffers by
 * elem_nrows*elem_ncols*1000 triples when full (_resize_buffers, :61-76). */
int feasm_coo_start(feasm_coo *coo, int64_t elem_nrows, int64_t elem_ncols, int64_t n_elem_mats,
                    int64_t row_nalldofs, int64_t col_nalldofs);
/* assemble! (src/AssemblyModule.jl:250-282): mat is nrows x ncols column-major;
 * rows/cols are 1-based; serialized column-major.  Errors exactly as the
 * reference: "Wrong size of matrix", "Column degree of freedom < 1", ... */
int feasm_coo_assemble(feasm_coo *coo, const double *mat, int64_t mat_nrows, int64_t mat_ncols,
                       const int64_t *rows, int64_t nrows, const int64_t *cols, int64_t ncols);
/* Many same-size element matrices at once: mats[e] column-major, rows/cols
 * nmats x nrows and nmats x ncols row-major. */
int feasm_coo_assemble_many(feasm_coo *coo, int64_t nmats, const double *mats, const int64_t *rows,
                            int64_t nrows, const int64_t *cols, int64_t ncols);
/* setnomatrixresult (src/AssemblyModule.jl:29-33). */
int feasm_coo_set_nomatrixresult(feasm_coo *coo, int flag);
/* makematrix! (src/AssemblyModule.jl:304-329): *out = NULL when
 * nomatrixresult is set (the reference returns spzeros and keeps the
 * triples); otherwise the device sparse() runs and the pointer resets to 1. */
int feasm_coo_makematrix(feasm_coo *coo, feasm_matrix **out);
/* number of triples currently buffered (buffer pointer - 1) */
int64_t feasm_coo_ntriples(const feasm_coo *coo);
int feasm_coo_destroy(feasm_coo *coo);
/* Device sparse(I, J, V, m, n) on device arrays (0-based int32 indices):
 * stable radix sort by (col, row), duplicates summed left to right in input
 * order, explicit zeros kept -- SparseArrays.sparse as called by makematrix!. */
int feasm_sparse_device(feasm_ctx *ctx, int64_t ntrip, const int32_t *d_I, const int32_t *d_J, const double *d_V,
                        int64_t m, int64_t n, feasm_matrix **out);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  feasm_sparse_device into caller arrays -- d_colptr n+1 int64
 * (0-based), d_rowval and d_nzval with room for ntrip entries -- with the
 * same kernels and results (bitwise); *nnz receives the entry count.  Every
 * temporary lives in one fixed device array per GPU (36 MiB, held until the
 * process ends): at most FEASM_NOALLOC_MAX_TRIPLES triples per call, more is
 * FEASM_ENOMEM.  No device memory is allocated or freed.  Calls that overlap
 * on one GPU share that array and race: serialise them. */
#define FEASM_NOALLOC_MAX_TRIPLES 1048576
int feasm_sparse_device_noalloc(feasm_ctx *ctx, int64_t ntrip, const int32_t *d_I, const int32_t *d_J,
                                const double *d_V, int64_t m, int64_t n, int64_t *d_colptr, int32_t *d_rowval,
                                double *d_nzval, int64_t *nnz);

/* sparse(I, J, V, m, n) on host triples (1-based int64 indices, as Julia's
 * SparseArrays.sparse takes them at makematrix!, src/AssemblyModule.jl:319-325):
 * uploads them and runs the device sparse(). */
int feasm_sparse(feasm_ctx *ctx, int64_t ntrip, const int64_t *I, const int64_t *J, const double *V, int64_t m,
                 int64_t n, feasm_matrix **out);

/* ---- mesh generators (src/MeshHexahedronModule.jl:39-112,
 *      src/MeshTetrahedronModule.jl:53-261) ------------------------------- */
/* H8blockx on the box [0,L]x[0,W]x[0,H] with nL x nW x nH cells: node i along
 * an axis at coordinate L*i/nL.  Writes 1-based conn (row-major nelem x 8)
 * and xyz (row-major nnodes x 3); pass NULL arrays to query the sizes. */
int feasm_gen_h8block(double L, double W, double H, int64_t nL, int64_t nW, int64_t nH, int64_t *nelem,
                      int64_t *nnodes, int64_t *conn, double *xyz);
/* T10block(..., orientation = :a): T4blockx + T4toT10 (first-encounter
 * mid-edge numbering).  amp_vertex / amp_mid: seeded perturbation amplitudes
 * applied to interior vertex nodes before the mid-edge nodes are placed and
 * to interior mid-edge nodes afterwards (0 = none). */
int feasm_gen_t10block(double L, double W, double H, int64_t nL, int64_t nW, int64_t nH, double amp_vertex,
                       double amp_mid, uint64_t seed, int64_t *nelem, int64_t *nnodes, int64_t *conn, double *xyz);
/* Seeded perturbation of nodes strictly inside [lo, hi] (row-major xyz):
 * component c of node n moves by u*amp, u in [-1,1) from SplitMix64 at
 * counter 3n+c+1 of stream `seed`. */
int feasm_gen_perturb(int64_t nnodes, double *xyz, const double lo[3], const double hi[3], double amp,
                      uint64_t seed);

/* ---- batched multi-mesh assembly (no reference counterpart: the
 *      reference assembles one mesh per bilform_* call) ------------------- */
/* Many independent meshes (each with its own form, coefficients and
 * resident pattern) assembled by one persistent fused launch per element
 * type/arithmetic group, optionally captured in a hipGraph. */
int feasm_batch_create(feasm_ctx *ctx, feasm_batch **out);
/* Add a mesh that already has its DOFs (node-major numbering); the form
 * arguments are those of feasm_bilform.  H8/GaussRule(3,2) and T10/TetRule(4). */
int feasm_batch_add(feasm_batch *b, feasm_mesh *mesh, int form, int rule, int order, const double *coeff,
                    int ncoeff, int precision);
/* Build patterns, tiles and launch tables; use_graph != 0 captures the
 * launches in a hipGraph replayed by feasm_batch_run.  max_nodes (owned
 * nodes per tile, <= 255), budget_kib (LDS per workgroup) and threads (256, 512
 * or 1024 per workgroup) tune the tiles; 0 picks the defaults. */
int feasm_batch_prepare(feasm_batch *b, int use_graph, int max_nodes, int budget_kib, int threads);
/* Enqueue one batched assembly on the context stream (asynchronous). */
int feasm_batch_run(feasm_batch *b);
/* Device time of the last feasm_batch_run (waits for it). */
int feasm_batch_last_ms(feasm_batch *b, double *ms);
/* Result matrix of entry i (owned by the batch: do not destroy). */
int feasm_batch_matrix(feasm_batch *b, int64_t i, feasm_matrix **out);
int feasm_batch_stats(feasm_batch *b, int64_t *nmeshes, int64_t *nelem, int64_t *nnz, int64_t *ntiles);
/* Autotune max_nodes/budget over a small grid (timing `reps` runs each),
 * leaving the batch prepared with the winner.  With db_path and key the
 * result is cached in (and first looked up from) a text database. */
int feasm_batch_autotune(feasm_batch *b, const char *db_path, const char *key, int reps, int *best_nodes,
                         int *best_budget_kib, int *best_threads, double *best_ms);
int feasm_batch_destroy(feasm_batch *b);

/* ---- assembly server (continuous batching; no reference counterpart) --- */
/* Worker threads (one HIP stream each) take queued requests as soon as a
 * worker and device memory are free; admission is by an HBM budget of
 * hbm_fraction of the free memory at creation, evicting least recently used
 * resident meshes that are not in use.  Results are bitwise independent of
 * what else runs. */
int feasm_server_create(int device, int workers, double hbm_fraction, feasm_server **out);
/* Queue the assembly of a new mesh (arrays are copied; conn 1-based row-major
 * nelem x npe, xyz row-major nnodes x 3, node-major DOFs with ndn per node).
 * keep_resident != 0 keeps mesh, pattern and tiles on the device for later
 * re-assembly under *mesh_id.  A request whose device-memory estimate exceeds
 * the server's whole budget is refused here with FEASM_ENOMEM. */
int feasm_server_submit_mesh(feasm_server *s, int elem_type, int64_t nelem, int npe, const int64_t *conn,
                             int64_t nnodes, const double *xyz_rowmajor, int ndn, int form, int rule, int order,
                             const double *coeff, int ncoeff, int precision, int keep_resident, int64_t *request_id,
                             int64_t *mesh_id);
/* The same without copying: conn and xyz_rowmajor are borrowed and must stay
 * valid and unchanged until feasm_server_wait reports the request finished;
 * the worker that takes the request converts and checks them (an entry of
 * conn outside 1..nnodes fails the request with FEASM_EINVAL), so the
 * submitting thread does no per-element work. */
int feasm_server_submit_mesh_borrowed(feasm_server *s, int elem_type, int64_t nelem, int npe, const int64_t *conn,
                                      int64_t nnodes, const double *xyz_rowmajor, int ndn, int form, int rule,
                                      int order, const double *coeff, int ncoeff, int precision, int keep_resident,
                                      int64_t *request_id, int64_t *mesh_id);
/* Queue a re-assembly of a resident mesh with new coefficients (pattern and
 * tiles reused: numeric phase only). */
int feasm_server_submit_reassembly(feasm_server *s, int64_t mesh_id, int form, int rule, int order,
                                   const double *coeff, int ncoeff, int precision, int64_t *request_id);
/* Wait up to timeout_s; *status 2 = done, < 0 = failed (feasm status);
 * latency and queueing time in ms from submission. */
int feasm_server_wait(feasm_server *s, int64_t request_id, double timeout_s, int *status, double *latency_ms,
                      double *queue_ms);
/* Cancel a request: a queued one is withdrawn at once; a running one finishes
 * its device work and is then reported cancelled (result dropped, mesh not
 * kept).  feasm_server_wait then gives status FEASM_ECANCELLED.  A request
 * that already ended is FEASM_ESTATE. */
int feasm_server_cancel(feasm_server *s, int64_t request_id);
/* Result of a finished request; owned by the server until released. */
int feasm_server_result(feasm_server *s, int64_t request_id, feasm_matrix **out);
int feasm_server_release(feasm_server *s, int64_t request_id);
int feasm_server_stats(feasm_server *s, int64_t *queued, int64_t *running, int64_t *completed, int64_t *residents,
                       int64_t *bytes_in_use, int64_t *budget);
const char *feasm_server_last_error(feasm_server *s);
int feasm_server_destroy(feasm_server *s);

#ifdef __cplusplus
}
#endif
#endif /* FEASM_H */
