/* cqgpu.h -- C ABI of the MI355X direct-SCF two-electron contraction G[D].
 *
 * Replaces, behind ChronusQ's own operator API, the body of
 *   template<typename Op, typename T>
 *   void AOIntegrals::newTwoEContractDirect(const std::vector<std::reference_wrapper<Op>>& X,
 *        std::vector<std::reference_wrapper<Op>>& AX,
 *        std::vector<ERI_CONTRACTION_TYPE>& contractionList, std::vector<T>& scalingFactors)
 * (include/aointegrals/aointegrals_contract.h:24-194, declared include/aointegrals.h:474-480),
 * together with the state it reads: the shell list (BasisSet::shells, mapSh2Bf, nShell;
 * include/basisset.h), the Schwarz bounds (AOIntegrals::computeSchwartz,
 * src/aointegrals/aointegrals_builders.cpp:469-518) and the screening threshold
 * (AOIntegrals::thresholdSchwartz_, include/aointegrals.h:334).
 *
 * Conventions (all as in the reference):
 *   - matrices are column-major N x N with leading dimension ld (Eigen default);
 *   - N counts spherical functions for pure shells (2l+1, m = -l..l) and Cartesian
 *     functions otherwise (x^l ... z^l, each carrying the x^l normalisation);
 *   - AX accumulates; callers zero it first (include/singleslater/singleslater_fock.h:47-67);
 *   - COULOMB adds 2*scale*J[X], EXCHANGE adds 2*scale*K[X]; scale == 0 matrices are not
 *     contracted but still enter the screening bound.
 * Errors are return codes (never exceptions); cq_last_error() gives the message.  Calls on one
 * context are synchronous (cq_contract_direct) and not thread-safe.
 */
#ifndef CQGPU_H
#define CQGPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cq_ctx cq_ctx_t;

/* return codes */
#define CQ_OK 0
#define CQ_ERR_INVALID -1   /* bad argument (reference: assert / CErr) */
#define CQ_ERR_NOMEM -2     /* device allocation failed (reference: std::bad_alloc) */
#define CQ_ERR_DEVICE -3    /* HIP runtime error */

/* ERI_CONTRACTION_TYPE (include/aointegrals.h:468-472); PAIR is not handled by the direct
 * path in the reference either. */
#define CQ_COULOMB 0
#define CQ_EXCHANGE 1

/* precision modes */
#define CQ_PREC_FP64 0
#define CQ_PREC_MIXED 1

/* Create a context from a flat shell list (replaces the BasisSet/AOIntegrals state that
 * newTwoEContractDirect reads, aointegrals_contract.h:45-48,96-109).
 *   l[s], pure[s], nprim[s]: angular momentum (0..2), pure flag, primitive count;
 *   alpha, coef: exponents and normalised contraction coefficients, concatenated by shell
 *               (the libint2::Shell coefficients after BasisSet::renormShells);
 *   origin: 3*nshell centre coordinates in bohr.
 * Shell order defines function order (mapSh2Bf, src/basisset/basisset_map.cpp:29-40).
 * Computes the Schwarz bounds on the device (computeSchwartz). */
int cq_ctx_create(cq_ctx_t** ctx, int nshell, const int* l, const int* pure, const int* nprim,
                  const double* alpha, const double* coef, const double* origin);
int cq_ctx_destroy(cq_ctx_t* ctx);

/* A context for several molecules at once (no reference counterpart; the reference runs one
 * molecule per job): the shells of all molecules concatenated, molecule[s] = 0, 1, ...
 * nondecreasing over the shells.  Shell pairs and quartets are formed only within a molecule,
 * so one contraction over the block-diagonal N x N density (N = all molecules' functions)
 * gives every molecule's own G[D] in its diagonal block, with one launch per class pair for
 * the whole batch.  cq_one_electron rejects such a context. */
int cq_ctx_create_batch(cq_ctx_t** ctx, int nshell, const int* l, const int* pure, const int* nprim,
                        const double* alpha, const double* coef, const double* origin, const int* molecule);

/* Message for the last failing call on ctx (or of cq_ctx_create when ctx is NULL). */
const char* cq_last_error(const cq_ctx_t* ctx);

/* Schwarz screening threshold tau; the reference hard-codes 1e-14
 * (include/aointegrals.h:334) and has no setter. */
int cq_set_schwarz_threshold(cq_ctx_t* ctx, double tau);
int cq_get_schwarz_threshold(const cq_ctx_t* ctx, double* tau);

/* Precision of the quartet arithmetic (no reference counterpart: the reference is FP64).
 * CQ_PREC_FP64: every quartet in FP64.  CQ_PREC_MIXED: quartets whose density-weighted
 * Schwarz bound (the screening bound max|X block| Q12 Q34) lies in [tau, threshold) are
 * evaluated with FP32 integrals (Boys function, J/K accumulation into G and the reduction
 * stay FP64); bounds >= threshold stay FP64.  Inside an incremental Fock build X is the
 * density difference, so the bound is Schwarz(ab) Schwarz(cd) max|dD|.  The FP32 kernels
 * serve one-density contractions on the LDS segment path (RHF G[D]); other calls run FP64
 * throughout.  threshold > 0 (1e-5 by default). */
int cq_set_precision_mode(cq_ctx_t* ctx, int mode, double threshold);

/* Sizes: shells, basis functions, Cartesian functions, shell pairs, primitive pairs kept. */
int cq_get_dims(const cq_ctx_t* ctx, int* nshell, int* nbf, int* ncart, int* npair, int* nprim_pair);

/* Q(s1,s2) = sqrt(max |(ij|ij)|), nshell x nshell symmetric (AOIntegrals::schwartz_). */
int cq_get_schwarz(const cq_ctx_t* ctx, double* Q);

/* newTwoEContractDirect on host matrices: X[i], AX[i] column-major n x n, leading
 * dimension ld; AX[i] += 1/2 scale[i] G_i (aointegrals_contract.h:189-193).  Synchronous. */
int cq_contract_direct(cq_ctx_t* ctx, int nmat, const double* const* X, double* const* AX,
                       const int* type, const double* scale, int n, int ld);

/* The same contraction on device-resident matrices, enqueued on stream (a hipStream_t;
 * NULL selects the context's stream).  Returns once the work is enqueued. */
int cq_contract_direct_device(cq_ctx_t* ctx, int nmat, const double* const* X, double* const* AX,
                              const int* type, const double* scale, int n, int ld, void* stream);

/* cq_contract_direct_device without any device allocation (no reference counterpart; for callers
 * that cannot allocate: code under HIP graph capture, tight serving loops).  The per-call
 * workspace -- 5 ncart^2 doubles per distinct density with a non-zero scale, plus one int each --
 * lives in a 64 MiB device array of the library, one per device for the life of the process; a
 * call that needs more returns CQ_ERR_INVALID.  Every call on a device shares that array: calls
 * that overlap on one device (concurrent streams or threads) race.  The context's own buffers
 * (task lists, counters) are allocated at cq_ctx_create / cq_set_task_budget, never here.
 * Profiling (cq_set_profiling) records no events on this path. */
int cq_contract_direct_device_noalloc(cq_ctx_t* ctx, int nmat, const double* const* X, double* const* AX,
                                      const int* type, const double* scale, int n, int ld, void* stream);

/* One-electron integrals S, T, V in the shells' basis, column-major n x n (leading dimension
 * ld) in device memory; Z[natom] and xyz[3*natom] (bohr) on the host.  Replaces
 * AOIntegrals::computeAOOneE (src/aointegrals/aointegrals_builders.cpp:337-467).  Synchronous. */
int cq_one_electron(cq_ctx_t* ctx, int natom, const double* Z, const double* xyz, double* S, double* T,
                    double* V, int ld, void* stream);

/* The same for a context of several molecules (cq_ctx_create_batch): atom_molecule[natom] gives
 * each nucleus's molecule (nondecreasing), and every molecule's shell pairs feel only its own
 * nuclei, so the outputs are the molecules' own S, T, V in their diagonal blocks.  No
 * reference counterpart. */
int cq_one_electron_batch(cq_ctx_t* ctx, int natom, const double* Z, const double* xyz, const int* atom_molecule,
                          double* S, double* T, double* V, int ld, void* stream);

/* Electric multipole integrals <mu| x^i y^j z^k |nu> about origin[3] (host, bohr; NULL: the
 * origin), 1 <= i + j + k <= order <= 3, in the shells' basis: 3, 9 or 19 column-major n x n
 * matrices (leading dimension ld) back to back in device memory `out`, in the order x, y, z;
 * xx, xy, xz, yy, yz, zz; xxx, xxy, xxz, xyy, xyz, xzz, yyy, yyz, yzz, zzz.  Replaces
 * AOIntegrals::OneEDriver with libint2::Operator::emultipole1..3 (origin 0; the reference keeps
 * the overlap as its first block: cq_one_electron gives it), src/aointegrals/
 * aointegrals_builders.cpp:38-55,387-392.  Synchronous. */
int cq_multipole(cq_ctx_t* ctx, int order, const double* origin, double* out, int ld, void* stream);

/* Real solid-harmonic coefficients for angular momentum l (0..2): (2l+1) x ncart(l) row-major,
 * m = -l..l, in the shell's uniformly normalised Cartesian functions (x^l ... z^l) -- the
 * transform of every pure shell (cart2sphCoeff, src/basisset/basisset_read.cpp:871-990). */
int cq_solid_harmonics(int l, double* out);

/* Surviving quartets per class pair of the last contraction (synchronises). */
int cq_get_quartet_counts(cq_ctx_t* ctx, long long* counts, int max_counts, int* n_out);
/* Surviving quartets of the last contraction by arithmetic: FP64 integrals and FP32 integrals
 * (the mixed-precision window; 0 in FP64 mode), summed over class pairs. */
int cq_get_precision_counts(cq_ctx_t* ctx, long long* fp64, long long* fp32);
int cq_num_class_pairs(void);
int cq_class_pair_info(int k, int* la, int* lb, int* lc, int* ld);

/* J/K digestion mode: 2 (default) groups the tasks of each class pair by the pair class with
 * fewer pairs (bra or ket) and accumulates that pair's K rows/columns and J block in LDS,
 * flushed once per work item; 1 always groups by the ket pair; 0 sends every update to HBM as
 * an FP64 atomic.  Results agree to rounding in every mode. */
int cq_set_lds_digestion(cq_ctx_t* ctx, int mode);

/* Digestion variant of one class pair k (0 .. cq_num_class_pairs()-1): 0, 1 or 2 as in
 * cq_set_lds_digestion (capped by the context's mode), -1 to follow the context's mode.  Lets
 * a tuner pick the fastest variant per class pair from cq_get_kernel_times; results agree to
 * rounding in every variant. */
int cq_set_class_digestion(cq_ctx_t* ctx, int k, int mode);
/* Chunks per work item of class pair k's group kernels (no reference counterpart): the ket (or
 * bra) Cartesian components of the high-angular-momentum class pairs are split into up to 36
 * chunks by register budget; a work item takes a task region and several of its chunks, spread
 * over the workgroup's wavefronts, and flushes its LDS accumulators once.  cpi >= 1: at most cpi
 * chunks per item (the largest divisor of the chunk count not above cpi); 0 restores the
 * built-in choice. */
int cq_set_class_chunk_group(cq_ctx_t* ctx, int k, int cpi);
int cq_get_class_digestion(const cq_ctx_t* ctx, int k, int* mode);

/* Stream (0..3) the kernels of class pair k are launched on: 0 is the stream the contraction
 * was called on, 1..3 the context's auxiliary streams; -1 (default) for round robin in launch
 * order (highest angular momentum first; the last, incomplete round from the last stream
 * down).  For a tuner that balances the streams by measured kernel times; no effect on
 * results. */
int cq_set_class_stream(cq_ctx_t* ctx, int k, int stream);

/* Number of HIP streams (1..4, default 4) the class-pair kernels of one contraction are spread
 * over, the stream the contraction was called on first; 1 runs them one after another on it
 * (for per-kernel timing).  Four streams use the process's four hardware queues one each. */
int cq_set_launch_streams(cq_ctx_t* ctx, int n);

/* Device memory for the screened task lists (no reference counterpart: the reference screens
 * and computes quartet by quartet on the host, aointegrals_contract.h:130-146).  The largest
 * task list of one class pair, in 8-byte entries (1 .. 2^31-1, default 2^28 = 2 GiB); a class
 * pair whose worst case (every quartet of its pairs passing the screen) is larger is screened
 * and digested in batches of its uniform groups, one after another, so memory stays bounded
 * whatever the system size.  Waits for the context's work and reallocates.  Results agree to
 * rounding for every budget. */
int cq_set_task_budget(cq_ctx_t* ctx, long long entries);
/* Task entries allocated over all class pairs, and the largest number of batches a class pair
 * takes. */
int cq_get_task_plan(const cq_ctx_t* ctx, long long* entries, int* max_batches);

/* Tile of class pair k's digestion kernels (no reference counterpart; set by the autotuner).
 * A work item digests one group of up to 16 uniform shell pairs sharing a shell against the
 * swept pairs that survive the screen against any of them; the J of each swept pair collects in
 * an LDS slot across the group (one HBM update per item instead of one per quartet).
 * lds_bytes: dynamic LDS budget of the kernel (0: 48 KiB); max_group: cap on the group size,
 * 1..16, within the budget (0: the class pair's built-in cap, which differs for one-molecule
 * and several-molecule contexts); union_bound: swept pairs per task region, 64..1024 (0: the
 * class pair's built-in bound).  Waits for the context's work and re-plans.  Results agree to
 * rounding for every tile. */
int cq_set_class_tile(cq_ctx_t* ctx, int k, long long lds_bytes, int max_group, int union_bound);
/* The planned tile for digestion variant 0/1 (ket-keyed) or 2 (bra-keyed): largest group, union
 * bound, number of groups, dynamic LDS bytes of the kernel. */
int cq_get_class_tile(const cq_ctx_t* ctx, int k, int variant, int* max_group, int* union_bound, int* ngroups,
                      long long* lds_bytes);

/* Profiling (no reference counterpart beyond its wall-clock G[P] timer,
 * include/singleslater/singleslater_fock.h:116-119): HIP events around every class-pair
 * kernel; accumulated milliseconds and launch counts per class pair. */
int cq_set_profiling(cq_ctx_t* ctx, int on);
int cq_get_kernel_times(cq_ctx_t* ctx, double* ms, long long* launches, int max_n);
/* The same for each class pair's screening kernel (the canonical quartet loop and the
 * density-weighted Schwarz test, aointegrals_contract.h:96-131); entry CQ_NUM_CLASS_PAIRS
 * (when max_n > CQ_NUM_CLASS_PAIRS) holds the shared launches that screen several class pairs
 * (cq_set_screen_fusion). */
int cq_get_screen_times(cq_ctx_t* ctx, double* ms, long long* launches, int max_n);

/* How the class pairs that screen once per call are screened: one launch per class pair on
 * its own stream (CQ_SCREEN_PER_CLASS), one launch per stream for the class pairs that run on
 * it, ahead of their kernels (CQ_SCREEN_PER_STREAM, the default), or one launch for all class
 * pairs before any class kernel starts (CQ_SCREEN_ONE).  The surviving quartets are identical
 * in every mode.  No reference counterpart (the reference screens inside its quartet loop,
 * aointegrals_contract.h:96-131). */
#define CQ_SCREEN_PER_CLASS 0
#define CQ_SCREEN_PER_STREAM 1
#define CQ_SCREEN_ONE 2
int cq_set_screen_fusion(cq_ctx_t* ctx, int mode);

/* Work sums of the last contraction per class pair, 4 doubles each: sum K_ab*K_cd
 * (primitive quartets), sum K_ab, Cartesian->spherical transform FLOPs, output integrals.
 * CQ_ERR_INVALID when a class pair runs in batches (cq_set_task_budget). */
int cq_get_work_stats(cq_ctx_t* ctx, double* out, int max_n);

/* Contracted quartets (s1 s2|s3 s4) in the shells' own basis, row-major per quartet,
 * concatenated (the buffer libint2::Engine::compute returns, aointegrals_contract.h:138-143). */
int cq_debug_quartets(cq_ctx_t* ctx, int nq, const int* shells4, double* out);

#ifdef __cplusplus
}
#endif
#endif /* CQGPU_H */
