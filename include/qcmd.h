/*
 * qcmd.h -- C-ABI of libqcmd.so, the MI355X (gfx950) McMurchie-Davidson
 * integral and Fock-build library.
 *
 * Conventions: every entry point returns 0 on success and a nonzero status
 * otherwise, with qc_last_error() describing the failure (the calling
 * thread's last error).  All array arguments are device pointers
 * (caller-owned, e.g. torch tensors' data_ptr()) unless documented as host;
 * `stream` is a hipStream_t (NULL = default stream).  Nothing here
 * allocates device memory except qc_init (per-device Boys tables).
 *
 * The reference (jax_qc, Python) has no FFI: its boundary is a set of Python
 * functions.  Each entry point below names the reference function it
 * replaces; the Python layer (jax_qc package in jax-examples_amd/) keeps the
 * reference's names and signatures on top of these.
 */
#ifndef QCMD_H
#define QCMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-device setup: builds the FP64/FP32 Boys tables on `device`.  Must be
 * called once per device before any other entry point on it. */
int qc_init(int device);
const char* qc_last_error(void);

/* F_0..F_nmax at each of n points: F[i*(nmax+1)+k] = F_k(t[i]).
 * Replaces jax_qc/integrals/boys.py:58-109 (boys_f0, boys_fn). */
int qc_boys(int nmax, const double* t, double* F, int64_t n, void* stream);
int qc_boys_f32(int nmax, const float* t, float* F, int64_t n, void* stream);

/* Primitive-pair table record (12 reals): p, P(3), P-A(3), P-B(3),
 * K = c_a c_b exp(-ab/p |A-B|^2), a (exponent on the first centre). */
#define QC_PP_STRIDE 12

/* Arguments of one ERI class launch (see qc_eri). */
typedef struct {
  const double* pairs64;  /* [npp][12] primitive-pair table, FP64 */
  const float* pairs32;   /* same in FP32 (FP32 launches only) */
  const int* sp_prim_off; /* per shell pair: first primitive pair */
  const int* sp_nprim;    /* per shell pair: primitive-pair count */
  const int* sp_ao;       /* per shell pair: (first AO of shell 1, of shell 2) */
  const int* bra_list;    /* shell-pair ids of the bra pair class */
  const int* ket_list;    /* shell-pair ids of the ket pair class */
  const int64_t* offsets; /* [nbra+1] prefix sum of surviving kets per bra */
  const int* kstart;      /* [nbra] first ket position per bra */
  int nbra;
  int64_t nquartets;
  int mode;               /* 0 dense n^4 tensor, 1 packed supermatrix, 2 Schwarz */
  int64_t nbf;
  double* out;            /* dense tensor or tile-major supermatrix (mode 0/1) */
  double* schwarz;        /* per shell pair bound (mode 2) */
  const double* boys64;   /* filled in by the library */
  const float* boys32;
  int variant;            /* 0 the per-class form measured faster on MI355X (one wave per quartet for (dp|dp), (dd|pp), (dd|dp), (dd|dd), one thread otherwise), 1 one thread per quartet, 2 one wave per quartet */
  const int64_t* sp_base; /* mode 1, optional: per shell pair, element offset of the pair's
                             supermatrix in out (many molecules in one launch: every
                             quartet's bra and ket pair belong to one molecule, whose AO
                             indices in sp_ao are molecule-local) */
  int block;              /* one-thread-per-quartet launches: quartets per work-group,
                             64 or 128 (0 = 128) */
} qc_eri_args;

/* One angular-momentum class of shell-quartet ERIs.  cab, ccd are pair-class
 * ids la(la+1)/2+lb with la>=lb, cab>=ccd, l<=2; fp32 selects FP32 primitive
 * sums.  Replaces jax_qc/integrals/eri.py:69-138 (compute_eri_tensor),
 * obara_saika.py:227-332/394-426 (primitive/contracted_eri_block) and, in
 * mode 2, screening.py:32-58 (compute_shell_pair_bounds). */
int qc_eri(int cab, int ccd, int fp32, const qc_eri_args* args, void* stream);

/* One-electron S, T, V (which = 0, 1, 2) for all shell pairs of pair class
 * cab, written (both triangles) into the dense n x n matrix M.
 * Replaces overlap.py:67-85, kinetic.py:42-59, nuclear.py:62-81.
 * which = 3, 4, 5: exact dipole integrals <mu| x_k - O_k |nu> about the
 * origin O = atom_xyz[0..2] (natom, atom_z unused); they replace the
 * approximate ones of jax_qc/properties/dipole.py:26-89. */
int qc_one_electron(int which, int cab, const double* pairs64, const int* sp_prim_off,
                    const int* sp_nprim, const int* sp_ao, const int* pair_list, int npairs,
                    int natom, const double* atom_xyz, const double* atom_z, int64_t nbf,
                    double* M, void* stream);

/* Supermatrix post-processing.  kind 0: P = J - 1/4[(ik|jl)+(il|jk)] (RHF
 * Raffenetti supermatrix); kind 1: K = 1/2[(ik|jl)+(il|jk)].  Input and
 * output are tile-major packed supermatrices over nbf functions. */
int qc_sm_assemble(int kind, const double* Jsm, double* out, int64_t nbf, void* stream);

/* qc_sm_assemble that also writes each tile's max |element| of `out` to
 * tmax[T(T+1)/2] (T = tiles per dimension): the tile maxima a compaction
 * needs (qc_sm_tile_maxabs), without another read of the supermatrix. */
int qc_sm_assemble_maxabs(int kind, const double* Jsm, double* out, int64_t nbf, double* tmax,
                          void* stream);

/* J <- J - K/2 in place over two tile-major supermatrices of nbf functions:
 * turns J into the P of kind 0 above once K (kind 1) is assembled, with no
 * third supermatrix.  J and K must be distinct.  tmax (may be NULL): each
 * tile's max |element| of the result, as qc_sm_tile_maxabs would give. */
int qc_sm_p_from_jk(double* J, const double* K, int64_t nbf, double* tmax, void* stream);

/* Launch shape of the Fock-build tile pass: tiles per work item (1..64) and
 * whether tile loads are non-temporal.  Process-wide. */
int qc_symv_config(int seg, int nontemporal);

/* Waves per work item of the tile pass for nrhs right-hand sides (1 or 2;
 * with 2 each wave takes half of every tile's rows, which halves the
 * registers per lane).  Process-wide; defaults 1 for nrhs = 1, 2 above. */
int qc_symv_split(int nrhs, int waves);

/* Workspace (bytes) for qc_symv over nbf functions and nrhs vectors. */
int64_t qc_symv_workspace_bytes(int64_t nbf, int nrhs);

/* y[r] = M x[r] for r < nrhs with M the symmetric tile-major supermatrix;
 * x, y are [nrhs][npair_pad] packed-pair vectors. */
int qc_symv(const double* M, int64_t nbf, int nrhs, const double* x, double* y, void* work,
            void* stream);

/* qc_symv without a caller workspace and without allocating: the row and
 * column partials live in a 64 MiB device array of the library (one per GPU,
 * its address taken at qc_init, held until the process ends).  No
 * counterpart in the reference; it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops).  Fails with a clear
 * error when the workspace of the configured tile-pass layout would exceed
 * the array (e.g. nbf <= 341 at nrhs = 1 and the default segment, 211 at
 * nrhs = 2) or when qc_init was not called on the current device.
 * Calls that overlap on one GPU share the array and race. */
int qc_symv_noalloc(const double* M, int64_t nbf, int nrhs, const double* x, double* y,
                    void* stream);

/* The two phases of qc_symv, callable separately (e.g. to time the
 * HBM-streaming tile pass alone): partial sums into `work`, then y. */
int qc_symv_tiles(const double* M, int64_t nbf, int nrhs, const double* x, void* work,
                  void* stream);
int qc_symv_reduce(int64_t nbf, int nrhs, double* y, void* work, void* stream);

/* Mixed-precision supermatrix: per-tile max |element| (tmax[T(T+1)/2]),
 * packing into a pool under a caller-built directory (fmt[t] = 0 skip,
 * 1 FP32 (16384 B), 2 FP64 (32768 B), 3 int16 (16 + 8192 B), 4 int8
 * (16 + 4096 B); off[t] = byte offset in the pool, a multiple of 16), and the
 * tile pass of the Fock build over that pool (accumulation in FP64;
 * qc_symv_reduce finishes).  An int16 / int8 tile is an FP64 scale
 * s = max|x| / 32767 (/ 127) and 8 bytes of padding, then the 64 x 64
 * integers rint(x / s) row-major; qc_sm_pack computes s itself. */
int qc_sm_tile_maxabs(const double* M, int64_t nbf, double* tmax, void* stream);
int qc_sm_pack(const double* M, int64_t nbf, const uint8_t* fmt, const int64_t* off, void* pool,
               void* stream);
int qc_symv_mixed_tiles(const void* pool, const uint8_t* fmt, const int64_t* off, int64_t nbf,
                        int nrhs, const double* x, void* work, void* stream);

/* Screened Fock build over a mixed pool, for builds from a density
 * difference (incremental SCF): the same contraction as the reference's
 * J/K einsums (jax_qc/scf/fock.py:31-38) with the tiles whose
 * tmax[t] * max(xmax[R], xmax[C]) < tau left out (each left-out tile moves
 * any element of y by less than 64 tau).  tmax: per-tile max |element| of
 * the pool (FP32); xmax[B]: max |x| over block B's 64 packed rows and every
 * right-hand side, from qc_block_maxabs; tau >= 0 (0 still skips zero
 * tiles).  qc_symv_screened_reduce finishes the build and must get the same
 * fmt, tmax, xmax and tau as the tile pass. */
int qc_block_maxabs(const double* x, int64_t nbf, int nrhs, float* xmax, void* stream);
int qc_symv_screened_tiles(const void* pool, const uint8_t* fmt, const int64_t* off,
                           const float* tmax, int64_t nbf, int nrhs, const double* x,
                           const float* xmax, double tau, void* work, void* stream);
int qc_symv_screened_reduce(int64_t nbf, int nrhs, double* y, void* work, const uint8_t* fmt,
                            const float* tmax, const float* xmax, double tau, void* stream);

/* x[r][pair(k,l)] = D_r[k][l] * (k == l ? 1 : 2) for r < nrhs. */
int qc_pack_density(const double* D, int64_t nbf, int nrhs, double* x, void* stream);

/* F_r = H + scale * unpack(y_r) (symmetric), r < nrhs (H may be NULL). */
int qc_unpack_fock(const double* H, const double* y, double scale, int64_t nbf, int nrhs,
                   double* F, void* stream);

/* ---- SCF step densities: trace-correcting purification (TC2) ---------- */
/* One TC2 step on nb symmetric n x n matrices X (contiguous, batch-major)
 * is qc_tc2_partials(X) -> X2 = X X (any GEMM) -> qc_tc2_update(X, X2):
 * X_b <- X2_b if tr X_b > nocc[b], else 2 X_b - X2_b, the branch taken on
 * the device; idem[b] = |tr X_b - tr X_b^2| and (if trace is not NULL)
 * trace[b] = tr X_b of the matrix before the step; a matrix whose
 * |tr X_b - tr X_b^2| is already below freeze is left unchanged (a
 * converged X stays converged however many further steps run; 0 disables
 * this).  part holds
 * qc_tc2_partials_len(nb) doubles of per-block partial sums.  No
 * allocation and no host synchronisation: a chunk of steps can be captured
 * in a graph.  Replaces the eigendecomposition of rhf.py:52-92 /
 * uhf.py:187-287 for the step densities (the SCF drivers still diagonalise
 * once at convergence for the orbitals). */
int64_t qc_tc2_partials_len(int nb);
int qc_tc2_partials(const double* X, int64_t n, int nb, double* part, void* stream);
int qc_tc2_update(double* X, const double* X2, const double* part, const double* nocc, int64_t n,
                  int nb, double* idem, double* trace, double freeze, void* stream);

/* The same TC2 step in one launch: Y_b = X_b^2 if tr X_b > nocc[b], else
 * 2 X_b - X_b^2 (X_b unchanged once |tr X_b - tr X_b^2| < freeze), with the
 * product computed on FP64 matrix cores over 32 x 32 tiles of the lower
 * triangle (X symmetric) and Y written in full.  part_in holds the per-tile
 * partial sums (tr, |X|_F^2) of X -- from qc_tc2_tile_partials or from the
 * previous step's part_out -- and part_out receives those of Y; both hold
 * qc_tc2_tile_partials_len(n, nb) doubles.  idem[b] = |tr X_b - tr X_b^2|.
 * X and Y, part_in and part_out must differ (steps ping-pong between two
 * buffers).  No allocation, no host synchronisation (graph-capturable). */
int64_t qc_tc2_tile_partials_len(int64_t n, int nb);
int qc_tc2_tile_partials(const double* X, int64_t n, int nb, double* part, void* stream);
int qc_tc2_step(const double* X, double* Y, const double* part_in, double* part_out,
                const double* nocc, int64_t n, int nb, double* idem, double freeze, void* stream);

/* ---- many molecules per launch (batched SCF) -------------------------- */
/* qc_one_electron for shell pairs of many molecules: pair sp writes its
 * block at M + pair_base[sp] (row stride nmax, i.e. a padded
 * [nmol][nmax][nmax] tensor) with molecule-local AO indices in sp_ao, and
 * (which = 2) sums the nuclei pair_atoms[2 sp] .. + pair_atoms[2 sp + 1] of
 * the concatenated atom arrays. */
int qc_one_electron_batch(int which, int cab, const double* pairs64, const int* sp_prim_off,
                          const int* sp_nprim, const int* sp_ao, const int* pair_list, int npairs,
                          const int* pair_atoms, const double* atom_xyz, const double* atom_z,
                          int64_t nmax, const int64_t* pair_base, double* M, void* stream);

/* Each molecule m owns a tile-major supermatrix of molT[m] tile rows; the
 * molecules' tiles are numbered consecutively (first global tile tile0[m]),
 * as are their packed-vector rows (first row row0[m], 64 entries per row).
 * A Fock build for all of them is pack -> tiles -> reduce -> unpack, four
 * launches that a caller captures into one graph per SCF iteration.  The
 * batch counterpart of jax_qc/scf/fock.py:41-47 applied to every molecule
 * of a list (the reference runs molecules one at a time). */

/* Per-tile max |element| and format-directed packing over ntiles
 * consecutive 64 x 64 FP64 tiles (any number of molecules). */
int qc_tiles_maxabs(const double* M, int64_t ntiles, double* tmax, void* stream);
int qc_tiles_pack(const double* M, int64_t ntiles, const uint8_t* fmt, const int64_t* off,
                  void* pool, void* stream);

/* Host: item0[m] = first work item of molecule m for segment length seg
 * (item0[nmol] = total, returned).  Host arrays. */
int64_t qc_batch_items(const int64_t* molT, int nmol, int seg, int64_t* item0);

/* Tile pass for all molecules in one persistent launch.  M: FP64 tiles
 * (fmt = NULL) or a compacted pool with per-tile fmt/off (global tile
 * numbering).  work: (nitems + ntiles) * nrhs * 64 doubles. */
int qc_batch_symv_tiles(const void* M, const uint8_t* fmt, const int64_t* off, int nmol,
                        const int64_t* molT, const int64_t* tile0, const int64_t* item0,
                        const int64_t* row0, int64_t nitems, int64_t ntiles, int64_t npad,
                        int nrhs, int seg, const double* x, void* work, void* stream);
/* y[h][row*64 + r] for all nrows = row0[nmol] packed rows. */
int qc_batch_symv_reduce(int nmol, const int64_t* molT, const int64_t* tile0, const int64_t* item0,
                         const int64_t* row0, int64_t nrows, int64_t nitems, int64_t npad, int nrhs,
                         int seg, double* y, void* work, void* stream);
/* D: [nrhs][nmol][nmax][nmax], molecule m in its leading nbf[m] block. */
int qc_batch_pack_density(const double* D, int64_t nmax, int nmol, const int64_t* nbf,
                          const int64_t* row0, int nrhs, double* x, int64_t npad, void* stream);
/* F[h][m] = H[m] + scale * unpack(y) inside the nbf[m] block, H[m] outside. */
int qc_batch_unpack_fock(const double* H, const double* y, double scale, int64_t nmax, int nmol,
                         const int64_t* nbf, const int64_t* row0, int nrhs, int64_t npad, double* F,
                         void* stream);

/* J_mn = sum_ls D_ls (mn|ls), K_mn = sum_ls D_ls (ms|ln) from a dense
 * tensor.  Replaces jax_qc/scf/fock.py:31-38 (_coulomb, _exchange). */
int qc_dense_jk(const double* eri, const double* D, int64_t nbf, double* J, double* K,
                void* stream);

/* ---- analytic nuclear gradient (contracted on the fly) ---------------- */

/* Arguments of one two-electron gradient class launch.  Densities are in
 * the Cartesian AO basis (ncart x ncart).  G_abcd = Dj_ab Dj_cd -
 * cx sum_s (Dk_s,ac Dk_s,bd + Dk_s,ad Dk_s,bc); RHF: Dj = Dk0 = D, nk = 1,
 * cx = 1/4; UHF: Dj = Da + Db, Dk = {Da, Db}, nk = 2, cx = 1/2. */
typedef struct {
  const double* pairs64;
  const int* sp_prim_off;
  const int* sp_nprim;
  const int* sp_aoc;   /* per shell pair: Cartesian AO offsets (shell 1, shell 2) */
  const int* sp_atom;  /* per shell pair: atom of shell 1, atom of shell 2 */
  const int* sp_diag;  /* per shell pair: 1 when both shells are one shell */
  const int* bra_list;
  const int* ket_list;
  const int64_t* offsets;
  const int* kstart;
  int nbra;
  int64_t nquartets;
  int64_t ncart;
  const double* dj;
  const double* dk0;
  const double* dk1;
  int nk;
  double cx;
  double* grad;        /* natom x 3, accumulated (atomic adds) */
  int variant;         /* 0 library default, 1 one thread per quartet (low
                          classes only), 2 one wave per quartet with the bra
                          side contracted with G first, 3 one wave per
                          quartet with one lane per primitive quartet
                          ((pp|pp), (ds|pp); in FP32 also (dp|ps), (dp|pp),
                          (dp|ds)), 4 form 2 with that contraction on the
                          matrix cores ((pp|pp), (ds|pp), (dp|ps), (dp|pp),
                          (dp|dp)), 5 one lane per primitive quartet with G
                          contracted with the bra E first, per bra primitive
                          in LDS ((pp|pp), (ds|pp), (dp|ps), (dp|pp), (dp|ds),
                          (dp|dp), (dd|ps), (dd|pp): the default for them); a
                          class without the form asked for returns an error */
  const double* schwarz; /* optional (with tau_density > 0): per shell pair Q */
  double tau_density;    /* skip a quartet when max|G_abcd| Q_ab Q_cd < tau_density */
  const float* pairs32;  /* fp32 = 1: FP32 primitive-pair table */
  int fp32;              /* 1: primitive sums in FP32 (either kernel form), each
                            quartet's contributions summed and accumulated in FP64 */
  int natom;             /* atoms in grad (0: not given).  With 1 <= natom <= 256 the
                            one-thread-per-quartet kernels sum a work-group's
                            contributions in LDS and add each atom's once */
} qc_grad2e_args;

/* 1/2 sum_quartets f sum_abcd G_abcd d(ab|cd)/dR for one class pair,
 * added into grad.  Replaces the dERI term of jax_qc/grad/rhf_grad.py:
 * 186-209 (which rebuilds all ERIs at R +- h e_k). */
int qc_grad2e(int cab, int ccd, const qc_grad2e_args* args, void* stream);

/* Launch-side density prescreen of one class launch's quartet list (bra i =
 * pair bra[i] with kets ket[kstart[i] .. kstart[i] + offsets[i+1] - offsets[i]]).
 * A quartet is kept when
 *   (Mj[a,b] Mj[c,d] + cx sum_s (Mk_s[a,c] Mk_s[b,d] + Mk_s[a,d] Mk_s[b,c]))
 *     * Q[bp] Q[kp] (1 + 1e-6) >= tau,
 * (a, b) = shells[2 bp], shells[2 bp + 1] (likewise c, d of kp); Mj, Mk0 and
 * the optional Mk1 are ns x ns row-major shell-block maxima of |Dj|, |Dk_s|.
 * That bound is at least the max |G_abcd| qc_grad2e's own density screen
 * computes, so the kept list gives the same gradient.  write_off == NULL:
 * count[i] = kept kets of bra i.  Otherwise bra i's kept kets go, in list
 * order, to out_ket[write_off[i] ...].  No counterpart in the reference (its
 * gradient has no screening): it lets the mixed path launch only the
 * quartets that survive (replacing a multi-pass torch formulation). */
int qc_grad2e_prescreen(int nbra, const int64_t* offsets, const int* bra, const int* ket,
                        const int* kstart, const int* shells, int ns, const double* Mj,
                        const double* Mk0, const double* Mk1, double cx, const double* Q,
                        double tau, int64_t* count, const int64_t* write_off, int* out_ket,
                        void* stream);

/* sum_ab (2 - delta_ij) [D_ab d(T+V)_ab/dR - W_ab dS_ab/dR] over the shell
 * pairs of class cab (including the nuclear-centre derivative of V), added
 * into grad.  Replaces the dH and dS terms of rhf_grad.py:186-209. */
int qc_grad1e(int cab, const double* pairs64, const int* sp_prim_off, const int* sp_nprim,
              const int* sp_aoc, const int* sp_atom, const int* sp_diag, const int* pair_list,
              int npairs, int natom, const double* atom_xyz, const double* atom_z,
              int64_t ncart, const double* D, const double* W, double* grad, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QCMD_H */
