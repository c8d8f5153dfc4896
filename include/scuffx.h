/*
 * scuffx.h -- C ABI of the MI355X BEM system path (assemble -> LU -> solve).
 *
 * Drop-in replacement, behind plain pointers and sizes, for the
 * reference entry points of the path:
 *   RWGGeometry::RWGGeometry         libs/libscuff/RWGGeometry.cc:353
 *   RWGGeometry::AssembleBEMMatrix   libs/libscuff/libscuff.h:364-365,
 *                                    AssembleBEMMatrix.cc:684-773
 *   RWGGeometry::AssembleRHSVector   libs/libscuff/AssembleRHSVector.cc:365-427
 *   HMatrix::LUFactorize             libs/libhmat/libhmat.h:311, LBWrappers.cc:151-170
 *   HMatrix::LUSolve                 libs/libhmat/libhmat.h:312-317, LBWrappers.cc:175-245
 *
 * Conventions
 *  - Matrices are column-major, complex numbers interleaved (re, im):
 *    element (r, c) of M is at M[r + c*ldm] as in HMatrix (HMatrix.cc:325-342).
 *  - Pointers named *_dev are HIP device pointers on the current device.
 *  - Every call runs on the stream set by scuffx_set_stream (default: the
 *    null stream) and returns after its work is complete.
 *  - Frequencies are in units of c / 1 micron (3e14 rad/s), lengths in microns
 *    (libs/libMatProp/libMatProp.cc:44).
 *  - Return value: 0 or a LAPACK-style info (> 0) on success paths that
 *    have one; negative SCUFFX_ERR_* on failure, with a message available
 *    from scuffx_last_error().
 */
#ifndef SCUFFX_H
#define SCUFFX_H

#ifdef __cplusplus
extern "C" {
#endif

#define SCUFFX_OK               0
#define SCUFFX_ERR_ARG         -1
#define SCUFFX_ERR_ALLOC       -2
#define SCUFFX_ERR_HIP         -3
#define SCUFFX_ERR_NUMERIC     -4
#define SCUFFX_ERR_UNSUPPORTED -5
#define SCUFFX_ERR_IO          -6
/* positive status of the refinements: the result is the best iterate found,
   but its residual is above the requested tolerance */
#define SCUFFX_NOT_CONVERGED    1

/* element types (the `prec` arguments): complex128 (LHM_COMPLEX), complex64,
   and real double (LHM_REAL: the matrices of imaginary frequencies,
   AllocateBEMMatrix(PureImagFreq=true), AssembleBEMMatrix.cc:778-783) */
#define SCUFFX_C128 64
#define SCUFFX_C64  32
#define SCUFFX_F64  8

typedef struct scuffx_geom scuffx_geom;

/* message of the last failed call on this thread */
const char *scuffx_last_error(void);
int scuffx_version(void);
/* releases every device buffer the library holds (workspaces, the device
   data of live geometries) while the HIP runtime is still up; afterwards
   the library makes no HIP call (destructors skip theirs) and compute calls
   fail.  Meant for process exit (the Python module registers it with
   atexit); no counterpart in the reference. */
int scuffx_shutdown(void);
/* HIP stream (hipStream_t) used by subsequent calls on this thread */
int scuffx_set_stream(void *hip_stream);

/* ---- geometry: RWGGeometry::RWGGeometry (RWGGeometry.cc:353-807) ---- */
/* .scuffgeo subset: OBJECT / MESHFILE / MATERIAL / ENDOBJECT */
int scuffx_geom_open(const char *scuffgeo_path, scuffx_geom **G);
int scuffx_geom_new(scuffx_geom **G);
/* material: "PEC", "CONST_EPS_<re>[+/-<im>i][_MU_..][_SIGMA_..]", "DRUDE_<wp>_<gamma>" */
int scuffx_geom_add_mesh(scuffx_geom *G, const double *V, int nV, const int *T, int nT,
                         const char *material);
int scuffx_geom_add_msh(scuffx_geom *G, const char *msh_path, const char *material);
int scuffx_geom_free(scuffx_geom *G);
int scuffx_geom_total_bfs(const scuffx_geom *G);
int scuffx_geom_num_surfaces(const scuffx_geom *G);
int scuffx_geom_counts(const scuffx_geom *G, int ns, int *nV, int *nP, int *nE);
/* first matrix row of surface ns and its number of basis functions
   (RWGGeometry::BFIndexOffset, RWGSurface::NumBFs, RWGGeometry.cc:708-724) */
int scuffx_geom_bf_range(const scuffx_geom *G, int ns, int *offset, int *nbf);
/* RWGGeometry::Mate[ns] (RWGGeometry.cc:733-770): the first earlier surface
   identical to ns (same mesh file, same exterior medium, both PEC or the same
   interior material), or -1; surfaces added from raw arrays have no mates */
int scuffx_geom_mate(const scuffx_geom *G, int ns);
/* eps and mu of region r (0: exterior) at omega (units of c / 1 um), as
   RWGGeometry::RegionMPs[r]->GetEpsMu (libMatProp.cc:501); eps2, mu2:
   re, im */
int scuffx_geom_epsmu(const scuffx_geom *G, int region, double omega_re, double omega_im, double *eps2, double *mu2);
/* per edge: iV1 iV2 iQP iQM iPPanel iMPanel PIndex MIndex Index (libscuff.h:120-146) */
int scuffx_geom_edges(const scuffx_geom *G, int ns, int *out9, double *lengths);
/* per panel: centroid[3] zhat[3] radius area (libscuff.h:100-112) */
int scuffx_geom_panels(const scuffx_geom *G, int ns, double *out8);
/* vertex coordinates of surface ns after its birth transform, 3 per vertex
   (RWGSurface::Vertices, libscuff.h:249) */
int scuffx_geom_vertices(const scuffx_geom *G, int ns, double *out3);

/* ---- geometrical transformations (.trans files) ----
   A scuffx_trans holds the list of GTComplex of a .trans file
   (ReadTransFile, GTransformation.cc:531-580): TRANSFORMATION tag ...
   ENDTRANSFORMATION sections and one-line TRANS entries; NULL or "" gives
   one identity entry tagged DEFAULT. */
typedef struct scuffx_trans scuffx_trans;
int scuffx_trans_open(const char *path, scuffx_trans **T);
int scuffx_trans_count(const scuffx_trans *T);
/* tag of entry i, NUL-terminated into buf[len] */
int scuffx_trans_tag(const scuffx_trans *T, int i, char *buf, int len);
int scuffx_trans_free(scuffx_trans *T);
/* surface label (OBJECT label, mesh file, or ByHand; RWGSurface::Label) */
int scuffx_geom_label(const scuffx_geom *G, int ns, char *buf, int len);
int scuffx_geom_set_label(scuffx_geom *G, int ns, const char *label);
/* RWGGeometry::CheckGTCList (RWGGeometry.cc:931-943): 0 if every
   transformation of every entry names an existing surface */
int scuffx_geom_check_trans(const scuffx_geom *G, const scuffx_trans *T);
/* RWGGeometry::Transform(GTCs[i]) (RWGGeometry.cc:888-912): moves the named
   surfaces (vertices, edge centroids; panels re-initialized) and sets
   SurfaceMoved; fails without moving anything if a label is unknown.
   The next assembly rebuilds the device geometry and near-pair lists; FIPPI
   records come from the process-wide cache where their keys recur. */
int scuffx_geom_transform(scuffx_geom *G, const scuffx_trans *T, int i);
/* the Transform of one entry naming surface ns, given as X' = R X + D
   (R row-major 3x3, NULL = identity; D NULL = 0); UnTransform reverts it */
int scuffx_geom_transform_surface(scuffx_geom *G, int ns, const double *R9, const double *D3);
/* RWGGeometry::UnTransform (RWGGeometry.cc:917-921) */
int scuffx_geom_untransform(scuffx_geom *G);
/* SurfaceMoved[ns] of the last scuffx_geom_transform (RWGGeometry.cc:774-781) */
int scuffx_geom_surface_moved(const scuffx_geom *G, int ns);
/* new coordinates for the nV vertices of surface ns, same mesh topology
   (the reference re-reads a mesh for that: RWGSurface.cc:119-217, then
   InitRWGPanel / InitEdgeList on the new vertices).  Panels and edges are
   re-derived, normals keep their orientation, the surface stops being a
   mate of another.  Identical coordinates are a no-op; an error if the
   geometry is transformed or a panel would degenerate. */
int scuffx_geom_set_vertices(scuffx_geom *G, int ns, const double *V, int nV);
/* precision of the far panel pairs in later assemblies of G: 64 (default,
   the reference's double-precision GetPPIs_Cubature,
   PanelPanelInteractions.cc:211-350) or 32 (the mixed form's option: the
   same 36-point moments in single precision, ~1e-7 relative error per far
   entry; near, self and singular terms stay double) */
int scuffx_geom_set_far_precision(scuffx_geom *G, int bits);

/* ---- adjoint contraction ----
   For a rigid motion of surface ns (displacement along x, y, z, then
   rotations about up to three axes through the origin given by their
   generators gamma[9 t .. 9 t + 8], as in scuffx_assemble_block):
       out[m] = sum_ij conj(x_i) (dM/dtheta_m)_ij y_j      (complex, re/im)
   over the whole matrix, i.e. the blocks (ns, b) and (b, ns), b != ns, of
   the GradM / dMdT path (AssembleBEMMatrixBlock, libscuff.h:500-505).
   With x the adjoint solution, -Re(out) is the gradient of a real function
   of the solution y of M y = b.  The derivative blocks are assembled one
   surface pair at a time and contracted on the GPU; the full dM is never
   stored.  x, y: device vectors of TotalBFs complex128. */
int scuffx_adjoint(scuffx_geom *G, int ns, double wr, double wi, const void *x, const void *y, int ntorque,
                   const double *gamma, double *out);

/* ---- vertex-position derivatives (no counterpart in the reference, whose
        derivatives are the rigid motions of GradM / dMdT; the panel-pair
        gradient structure follows PanelPanelInteractions.cc:132-199) ----
   Gradient arrays have one slot per vertex coordinate, surfaces in order:
   slot 3 (sum of NumVertices of earlier surfaces + v) + c;
   scuffx_vertex_slots gives their number. */
int scuffx_vertex_slots(scuffx_geom *G);
/* grad[theta] = sum_ij C_ij dM_ij / dtheta (complex128), with C = u v^T
   (C NULL; u, v device vectors of TotalBFs) or a dense column-major C
   (device, leading dimension ldc).  M: the matrix assembled at this
   frequency (device, column-major, ldm), used for the edge-length terms.
   Far panel pairs are differentiated analytically with the assembly's own
   rule, separated near pairs with the 78-point rule on each panel quarter, touching and self
   pairs by central differences (step 1e-4 of the panel radius) of their
   Taylor-Duffy integrals; nothing of dM is stored.  Touching surfaces
   (panels of two surfaces sharing a vertex) are SCUFFX_ERR_UNSUPPORTED. */
int scuffx_vertex_gradient(scuffx_geom *G, double wr, double wi, const void *M, long ldm, const void *u,
                           const void *v, const void *C, long ldc, void *grad);
/* grad[theta] = sum_i mu_i d b_i / dtheta for the plane-wave RHS of
   scuffx_rhs_planewave (same E0, nhat); mu, grad: device complex128 */
int scuffx_rhs_vertex_gradient(scuffx_geom *G, double wr, double wi, const double *E0, const double *nhat,
                               const void *mu, void *grad);
/* absorbed power of scuffx_power (crossed overlaps) for the currents KN:
   grad[theta] = its derivative at fixed KN (real, device doubles) and
   gamma (device complex128, TotalBFs) with dPAbs = Re(gamma^T dKN) */
int scuffx_pabs_vertex_gradient(scuffx_geom *G, const void *KN, double *grad, void *gamma);

/* ---- batched frequency sweep over several geometries of uneven order
        (no counterpart in the reference: the scattering caller's loop,
        scuff-scatter.cc:373-567, over many instances at once) ----
   For each of the ngeom geometries, the nfreq frequencies (complex, 2
   doubles each) are assembled, factored in complex64 by one batched LU
   (tournament pivoting, block width autotuned per order and batch size),
   solved for the plane wave (E0: 3 complex, nhat) and refined in complex128
   with refine_steps fixed steps, and their power computed -- all recorded
   into one hipGraph and launched once.  Instances left above a 1e-11
   relative residual, or with a factorization problem, are solved again by
   the step-by-step mixed path.  out[3 (g nfreq + f) + {0,1,2}] = PAbs, PExt,
   PScat; relres (optional) the final residuals; info (optional, 2 ints):
   instances solved again, LU block width of the first geometry; ms
   (optional, 5 doubles): host milliseconds of set-up, capture,
   instantiation, graph run, results.  The device buffers are kept for the
   next sweep of the same shapes (until scuffx_shutdown).  Returns
   SCUFFX_ERR_UNSUPPORTED if a frequency needs the short-wavelength kernels
   (their job lists are built on the host). */
int scuffx_sweep_batched(scuffx_geom **Gs, int ngeom, const double *omegas, int nfreq, const double *E0,
                         const double *nhat, int refine_steps, double *out, double *relres, int *info,
                         double *ms);
/* ---- frequency sweep ----
   The scattering caller's loop over frequencies on one geometry
   (scuff-scatter.cc:373-567, no transformations): per frequency the BEM
   matrix, its factorization (prec 64: complex128 LU; prec 32: complex64
   LU + complex128 refinement to 1e-13), the plane-wave right-hand side
   (E0: 3 complex as 6 doubles, nhat: 3), the solve and the power closure.
   omegas: count complex frequencies as re, im pairs; out[3k .. 3k+2] =
   PAbs, PExt, PScat of frequency k.  One stream, buffers allocated once
   per call; the FIPPI table is built at the first frequency. */
int scuffx_batch_sweep(scuffx_geom *G, const double *omegas, int count, const double *E0, const double *nhat,
                       int prec, double *out);

/* ---- HDF5 files (libs/libhmat/HDF5IO.cc) ----
   HDF5 is loaded at run time (libhdf5.so + libhdf5_hl.so from the loader path,
   $SCUFFX_HDF5_DIR or /opt/conda/lib); without it these return
   SCUFFX_ERR_IO.  Data are host arrays, column-major, complex interleaved.
   HMatrix::OpenHDF5Context / CloseHDF5Context (HDF5IO.cc:120-160): */
int scuffx_hdf5_open(const char *path, void **ctx);
int scuffx_hdf5_close(void *ctx);
/* HMatrix::ExportToHDF5 (HDF5IO.cc:171-236): dims {nc, nr} or {nc, 2 nr},
   int attributes Storage_Type = 0 and RealComplex */
int scuffx_hdf5_write_matrix(void *ctx, const char *name, const double *data, int nr, int nc, int is_complex);
/* HVector::ExportToHDF5 (HDF5IO.cc:436-475): dims {n} or {2 n}, RealComplex */
int scuffx_hdf5_write_vector(void *ctx, const char *name, const double *data, int n, int is_complex);
/* HMatrix(FileName, LHM_HDF5, Name) / HVector(...) (HDF5IO.cc:252-418,
   :497-540): call with data = NULL for the sizes, then with a buffer */
int scuffx_hdf5_read_matrix(const char *path, const char *name, double *data, int *nr, int *nc, int *is_complex);
int scuffx_hdf5_read_vector(const char *path, const char *name, double *data, int *n, int *is_complex);
/* test-mesh generator: V[3*(10*4^L+2)], T[3*20*4^L] */
int scuffx_icosphere(int level, double radius, const double *center, double *V, int *T);

/* ---- frequency-independent panel-pair integrals: FIPPICache
        (FIPPICache.cc:135-187), built once per geometry on the GPU ---- */
int scuffx_fippi_build(scuffx_geom *G);
int scuffx_fippi_count(const scuffx_geom *G);
/* host copy, 36 doubles per record (QIFIPPIData, libscuffInternals.h:282-292) */
int scuffx_fippi_records(scuffx_geom *G, double *out);
/* ---- PreloadCache / StoreCache (libscuff.h:662-663, FIPPICache.cc:189-414) ----
   process-wide FIPPI record cache: every table built later takes the records
   whose 15-float key it finds there, and adds the ones it computes.  File
   format of the reference: "FIPPICACHE\0", then 352-byte records (15 float32
   key, 4 bytes padding, 36 doubles).  Preload returns the number of records
   read (an existing key keeps its record); store returns the number written
   (0, file untouched, when the cache was preloaded from that same file and has
   not grown, as the reference does). */
long scuffx_preload_cache(const char *path);
long scuffx_store_cache(const char *path);
long scuffx_cache_size(void);
/* empties the process-wide cache (no counterpart in the reference, whose
   global cache lives as long as the process); tables already built keep
   their records */
int scuffx_clear_cache(void);

/* ---- RWGGeometry::AssembleBEMMatrix(Omega, M) (AssembleBEMMatrix.cc:684) ----
   fills all N x N entries of the complex128 matrix at M_dev */
int scuffx_assemble(scuffx_geom *G, double omega_re, double omega_im, void *M_dev, long ldm);
/* ---- AllocateBEMMatrix(PureImagFreq=true, Packed) + AssembleBEMMatrix
        (AssembleBEMMatrix.cc:778-783, HMatrix.cc:419-435) ----
   the real matrix of the reference (LHM_REAL): the real part of every entry,
   as AddEntry(cdouble) on a real HMatrix keeps; at an imaginary frequency of
   a compact geometry the imaginary parts vanish.  packed = 0: column-major
   M[r + c*ldm]; packed = 1: LHM_SYMMETRIC packed upper storage
   M[r + c*(c+1)/2] for r <= c (N(N+1)/2 doubles, ldm unused). */
int scuffx_assemble_real(scuffx_geom *G, double omega_re, double omega_im, double *M_dev, long ldm, int packed);
/* the real counterpart of scuffx_assemble_block (the T, U and dU blocks of
   the Casimir caller, CreateSC3Data.cc:100-157): real parts of the block and
   of its derivatives; packed = 1 (diagonal blocks without derivatives only)
   stores the block in packed upper storage, ldb unused */
int scuffx_assemble_block_real(scuffx_geom *G, int nsa, int nsb, double omega_re, double omega_im, double *B_dev,
                               long ldb, int packed, double *GradB_dev, long ldg, long gstride, int ntorque,
                               const double *gamma, double *dBdT_dev);
/* full symmetric n x n matrix A (column-major, lda) from packed upper
   storage P (HMatrix LHM_SYMMETRIC layout, HMatrix.cc:330-333); prec 64, 32, 8 */
int scuffx_unpack_symmetric(const void *P_dev, long n, void *A_dev, long lda, int prec);

/* ---- RWGGeometry::AssembleBEMMatrixBlock(nsa, nsb, Omega, 0, M, GradM)
        (libscuff.h:500-505, AssembleBEMMatrix.cc:349-391) ----
   block of surfaces nsa (rows) and nsb (columns): B (NBF_a x NBF_b,
   column-major, leading dimension ldb); nsa == nsb gives the diagonal block
   (no derivatives).  For nsa != nsb, if GradB_dev is not NULL, its derivatives with respect
   to a rigid displacement of surface nsa along x, y, z: three matrices at
   GradB_dev + m*gstride (complex entries), leading dimension ldg; with
   ntorque = 1..3 also the derivatives with respect to a rotation of surface
   nsa about each axis (dMdT; gamma = 9*ntorque doubles, the rotation
   generators Gamma[i + 3 j] of CreateGammaMatrix, EdgeEdgeInteractions.cc:241)
   at dBdT_dev + t*gstride (SurfaceSurfaceInteractions.cc:212-327,
   PanelPanelInteractions.cc:132-199) */
int scuffx_assemble_block(scuffx_geom *G, int nsa, int nsb, double omega_re, double omega_im,
                          void *B_dev, long ldb, void *GradB_dev, long ldg, long gstride,
                          int ntorque, const double *gamma, void *dBdT_dev);
/* timings of the last assembly in ms: fippi self far near stamp mirror;
   counts: far pairs, near pairs, self jobs, fippi records */
int scuffx_assemble_stats(const scuffx_geom *G, double *ms6, long *counts4);

/* ---- HMatrix::LUFactorize (LBWrappers.cc:151-170) ----
   prec 64: complex128 (zgetrf), prec 32: complex64 (cgetrf), prec 8: real
   double (dgetrf; pivots by |a|, the first maximum).  ipiv_dev: int[n],
   1-based row interchanges.  Returns LAPACK info.  (Packed symmetric
   matrices are factored by the caller as a full copy, scuffx_unpack_symmetric,
   with partial pivoting -- not with dsptrf/zsptrf's Bunch-Kaufman pivoting.) */
int scuffx_lu_factor(void *A_dev, long n, long lda, int *ipiv_dev, int prec);
/* ---- HMatrix::LUSolve (LBWrappers.cc:175-245); trans 'N', 'T' or 'C';
        prec 64, 32 or 8 ---- */
int scuffx_lu_solve(const void *LU_dev, long n, long lda, const int *ipiv_dev, char trans,
                    long nrhs, void *B_dev, long ldb, int prec);

/* ---- one-right-hand-side solve that allocates nothing.  No counterpart in
        the reference: it is for callers that cannot allocate device memory
        (code under HIP graph capture, tight serving loops, harnesses that
        forbid hipMalloc).  Same result as scuffx_lu_solve(trans 'N', nrhs 1)
        on the same factors.  prec 64 or 32; n <= 16384 (larger n:
        SCUFFX_ERR_ARG).  Its buffers live in one fixed device array per GPU
        (scuffx_lu_solve_noalloc_bytes() bytes, held until the process
        ends), so calls that overlap on one GPU race. ---- */
int scuffx_lu_solve_noalloc(const void *LU_dev, long n, long lda, const int *ipiv_dev, void *B_dev, int prec);
long scuffx_lu_solve_noalloc_bytes(void);

/* ---- HMatrix::LUInvert (LBWrappers.cc:251-335): A <- A^{-1} from the
        factors of scuffx_lu_factor (zgetri / dgetri); prec 64, 32 or 8.
   Returns the first zero diagonal entry of U (1-based; A unchanged) or 0. */
int scuffx_lu_invert(void *A_dev, long n, long lda, const int *ipiv_dev, int prec);

/* ---- quantities the Casimir caller reads from the factors
        (applications/scuff-cas3D/CasimirIntegrand.cc:51-146) ----
   log |det A| and arg det A of an LU-factored matrix (diagonal of U and the
   parity of the interchanges) */
int scuffx_lu_logdet(const void *LU_dev, long n, long lda, const int *ipiv_dev, int prec,
                     double *logabs, double *phase);
/* sum of the first ncols diagonal entries of A^{-1} D with the factors of A;
   D (n x ncols, ncols <= n, device) is overwritten by A^{-1} D
   (GetTraceMInvdM, CasimirIntegrand.cc:100-146, where D is N x N1) */
int scuffx_trace_solve(const void *LU_dev, long n, long lda, const int *ipiv_dev, void *D_dev, long ldd,
                       long ncols, int prec, double *tr_re, double *tr_im);

/* ---- mixed precision: complex64 LU of a complex128 matrix plus
        complex128 iterative refinement (no counterpart in the reference;
        the plain path above is the drop-in) ---- */
/* LU32 = complex64 copy of M, factored in place: P M = L U with L unit
   lower, ipiv 1-based row interchanges in LAPACK layout.  The pivots come
   from tournament pivoting (communication-avoiding LU: each 32-column panel
   picks its 32 pivot rows by a tree of partial-pivoting eliminations over
   256-row groups), which differs from row-by-row partial pivoting; the
   environment variable SCUFFX_MIXED_PIVOTING=partial selects row-by-row
   partial pivoting instead.  The block-row triangular solves use the
   explicit inverses of the 32x32 unit-lower panel blocks, and the trailing
   updates three real products per complex product (Gauss's 3M; the
   environment variable SCUFFX_GEMM_3M=0 selects four), so the factors
   carry larger rounding than scuffx_lu_factor(prec=32): they are the
   refinement's preconditioner.  Returns the first zero pivot or 0. */
int scuffx_lu32_from_c128(const void *M_dev, long n, long ldm, void *LU32_dev, long ld32,
                          int *ipiv_dev);
/* x = M^{-1} b (complex128) from the complex64 factors: x0 = LU32^{-1} b,
   then x += LU32^{-1}(b - M x) with the residual in complex128, at most
   maxiter corrections, stopping when |b - M x|_inf <= tol |b|_inf or when a
   correction does not lower the residual (that correction is undone).  If
   the tolerance is still not met, GMRES-IR (below) continues from the best
   iterate.  Returns 0 when the tolerance is met, SCUFFX_NOT_CONVERGED when
   it is not (x is the best iterate, *relres its residual), or a negative
   error. */
int scuffx_refine(const void *M_dev, long n, long ldm, const void *LU32_dev, long ld32,
                  const int *ipiv_dev, const void *b_dev, void *x_dev, double tol, int maxiter,
                  int *iters, double *relres);
/* GMRES-IR: x0 = LU32^{-1} b, then corrections d from restarted flexible
   GMRES(20) on M d = b - M x in complex128, right-preconditioned by the
   complex64 factors; maxiter bounds the total GMRES steps (*iters).  Same
   tolerance and returns as scuffx_refine. */
int scuffx_gmres_ir(const void *M_dev, long n, long ldm, const void *LU32_dev, long ld32,
                    const int *ipiv_dev, const void *b_dev, void *x_dev, double tol, int maxiter,
                    int *iters, double *relres);

/* ---- callers either side of the path ---- */
/* AssembleRHSVector for one plane wave in the exterior medium
   (AssembleRHSVector.cc:365-427, PlaneWave.cc:71-87); E0 = 3 complex
   as 6 doubles, nhat = propagation direction */
int scuffx_rhs_planewave(scuffx_geom *G, double omega_re, double omega_im, const double *E0,
                         const double *nhat, void *rhs_dev);
/* absorbed power (crossed overlaps, OPFT.cc:310-364) and extinction
   power from the solution KN and the RHS */
int scuffx_power(scuffx_geom *G, double omega_re, double omega_im, const void *KN_dev,
                 const void *rhs_dev, double *PAbs, double *PExt);

/* ---- energy/momentum-transfer power (GetEMTPFTMatrix, EMTPFT.cc:468-803, the
        reference's default power method, PFTOptions.h:84) ----
   Mext_dev: the exterior region's matrix from scuffx_assemble_region(G, w, 0,
   ...), column-major complex128.  PScat from the quadratic form of the
   solution KN with it (integral method EHDERIVATIVES2, EMTPFT.cc:334-362),
   PAbs = extinction (scuffx_power) - PScat.  Every surface must live in the
   exterior region (else SCUFFX_ERR_UNSUPPORTED). */
int scuffx_power_emt(scuffx_geom *G, double omega_re, double omega_im, const void *KN_dev, const void *rhs_dev,
                     const void *Mext_dev, long ldm, double *PAbs, double *PScat);
/* the BEM matrix with the contributions of one region only (0: exterior),
   complex128: the G and C integrals the EMT power reads
   (GetGCMatrixElements for that region, EMTPFT.cc:336-353) */
int scuffx_assemble_region(scuffx_geom *G, double omega_re, double omega_im, int region, void *M_dev, long ldm);

/* ---- measurement (no counterpart in the reference) ----
   on != 0: every launch of the hot kernels is bracketed by HIP events on
   its stream; totals are reset.  Kernel names: gemm_c128 gemm_c64
   panel_c128 panel_c64 trsm_c128 trsm_c64 far near self stamp fippi resid grad.
   work = algorithmic flops (gemm, trsm, far, near) or bytes (stamp, resid). */
int scuffx_profile(int on);
int scuffx_profile_read(const char *kernel, double *ms, double *work, long *launches);

#ifdef __cplusplus
}
#endif

#endif
