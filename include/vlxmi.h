/* vlxmi C ABI: hand-written entry points (the per-class J/K launchers and the
 * Schwarz kernels are declared in vlxmi_kernels.h).  All pointers are device
 * pointers unless stated; every call returns 0 on success, a negative
 * vlxmi status (vlxmi_status.h) for argument errors, or a positive hipError_t. */
#ifndef VLXMI_H
#define VLXMI_H
#include <stddef.h>
#include <stdint.h>
#include "vlxmi_kernels.h"
#ifdef __cplusplus
extern "C" {
#endif

/* Library ABI version (1). */
int vlxmi_version(void);

/* Shape of the Boys tables consumed as boys_func_table / boys_func_ft by every
 * kernel (the reference passes VeloxChem's tables under these names,
 * gen_coulomb.py:91-92; their layout is this library's own). */
int vlxmi_boys_table_size(uint32_t* nrow, uint32_t* ncol, uint32_t* nft);

/* Fills HOST buffers table[nrow*ncol] and ft[nft]; the caller uploads them. */
int vlxmi_boys_table_fill(double* table, double* ft);

/* out[t*(L+1)+m] = F_m(T[t]), the device Boys function (test probe). */
int vlxmi_boys_eval(const double* T, uint32_t n, uint32_t L, const double* table, double* out, void* stream);

/* As vlxmi_boys_eval through the table's coarse rows copied to LDS, the
 * evaluation some light-class kernels use (L = 0..4; test probe). */
int vlxmi_boys_eval_coarse(const double* T, uint32_t n, uint32_t L, const double* table, double* out,
                           void* stream);

/* Human-readable text of a status code (static string). */
const char* vlxmi_status_string(int code);

/* Adds to *out the number of primitive shell quartets of one exchange class that
 * pass QKmax_IJ * QKmax_KL * maxD_bd > thr (thr_over_maxD = thr / maxD_bd). */
int vlxmi_count_k_quartets(const uint32_t* ikI, const uint32_t* ikK, uint32_t nblk,
                           const double* qa, const uint32_t* qa_displ, const uint32_t* qa_cnt,
                           const double* qc, const uint32_t* qc_displ, const uint32_t* qc_cnt,
                           double thr_over_maxD, unsigned long long* out, void* stream);

/* Deterministic AO assembly (segmented gather-sum, no atomics):
 * out[seg_out[s]] = sum_{t = seg_off[s]}^{seg_off[s+1]-1} values[perm[t]], in list order.
 * Replaces the AO assembly of VeloxChem's host driver (not in the reference tree). */
int vlxmi_segment_sum(const double* values, const uint32_t* perm, const uint32_t* seg_off,
                      const uint32_t* seg_out, uint32_t nseg, double* out, void* stream);

/* Two-electron part of the Fock matrix in one deterministic pass (no atomics):
 * out[seg_out[s]] = sum_{t = seg_off[s]}^{seg_off[s+1]-1} (src[t] ? cK : cJ) * values[perm[t]],
 * values = the per-pair mat_J outputs followed by the mat_K outputs, src[t] = 0
 * for a Coulomb and 1 for an exchange contribution; with cJ = 2, cK = -1 (D
 * without the spin factor 2) this is the closed-shell G = 2J - K.  Replaces the
 * Fock combination of VeloxChem's host driver (not in the reference tree). */
int vlxmi_fock_assemble(const double* values, const uint32_t* perm, const uint8_t* src, const uint32_t* seg_off,
                        const uint32_t* seg_out, uint32_t nseg, double cJ, double cK, double* out, void* stream);

/* p[0..n) = 0 by a kernel (safe inside HIP-graph capture). */
int vlxmi_fill_zero(double* p, uint64_t n, void* stream);

/* Ket-split partial sums of the Coulomb launchers (no counterpart in the
 * reference, whose J kernels give each bra entry one writer,
 * gen_coulomb.py:285-297).  A Coulomb launch with ket_split > 1 stores chunk y's
 * partial sums at part[y * count_local + e] in a workspace the library keeps per
 * (current device, stream) -- launches on one stream share it, other streams get
 * their own -- and then calls vlxmi_ket_split_reduce on the same stream.
 * vlxmi_ket_workspace returns (in *out) a buffer of at least ndoubles doubles,
 * growing it after synchronizing the stream; vlxmi_release_workspaces frees all
 * of them after a device synchronization. */
int vlxmi_ket_workspace(size_t ndoubles, void* stream, double** out);
int vlxmi_release_workspaces(void);

/* out[e] += sum_{y < nsplit} part[y * count + e] for e < count, the chunks added
 * in order with compensated (Neumaier) summation: deterministic, about one
 * rounding of the total. */
int vlxmi_ket_split_reduce(const double* part, uint32_t nsplit, uint32_t count, double* out, void* stream);

/* Reference list order -> the block order of these kernels (host arrays in and
 * out; no counterpart in the reference, whose kernels take component-entry
 * lists: J pair lists with i <= j for one type and the ket list sorted by
 * |Q_kl D_kl| per entry, gen_coulomb.py:205-212,276-281; exchange CSR segments
 * sorted per component by Q_K, gen_exchange.py:189-232; any (i,k) order).
 * Components a reference list does not carry become fillers (zero Schwarz
 * factor, pair data and density: they add nothing).  Capacities are in
 * entries; a call with too small a capacity (0 to query) returns
 * VLXMI_ERR_SIZE with the entries needed in *out_count / *out_nnz.
 * vlxmi_j_list_to_blocks: mat_D == NULL converts a bra list (blocks by (I, J)),
 * otherwise a ket list (blocks by non-increasing max |Q D|); out_ref_index
 * (may be NULL) gives each block entry's reference entry, -1 for fillers, for
 * vlxmi_scatter_add.  Lists with i > j (one type) return VLXMI_ERR_LAYOUT. */
int vlxmi_j_list_to_blocks(uint32_t la, uint32_t lb, const uint32_t* first, const uint32_t* second,
                           const double* pair_data, const double* mat_Q, const double* mat_D, uint32_t count,
                           uint32_t capacity, uint32_t* out_first, uint32_t* out_second, double* out_pair_data,
                           double* out_mat_Q, double* out_mat_D, int64_t* out_ref_index, uint32_t* out_count);
int vlxmi_ik_list_to_blocks(uint32_t la, uint32_t lc, const uint32_t* ik_i, const uint32_t* ik_k, uint32_t count,
                            uint32_t capacity, uint32_t* out_i, uint32_t* out_k, int64_t* out_ref_index,
                            uint32_t* out_count);
int vlxmi_k_csr_to_blocks(uint32_t la, uint32_t lb, uint32_t count_a, const uint32_t* displs, const uint32_t* counts,
                          const double* Q, const uint32_t* D_inds, const double* pair_data, uint64_t capacity,
                          uint32_t* out_displs, uint32_t* out_counts, double* out_Q, uint32_t* out_D_inds,
                          double* out_pair_data, uint64_t* out_nnz);

/* dst[ref_index[e]] += src[e] for e < n with ref_index[e] >= 0 (device arrays):
 * block-order mat_J / mat_K back to the caller's reference-order entries. */
int vlxmi_scatter_add(const double* src, const int64_t* ref_index, uint64_t n, double* dst, void* stream);

/* Device checks of the block contract (synchronous: they wait for the stream
 * and return VLXMI_OK or VLXMI_ERR_LAYOUT).  J list: whole blocks, I <= J for
 * one type, and for a ket list (mat_D != NULL) D = 0 below the diagonal of
 * diagonal blocks and non-increasing block bounds max |Q D|.  (i,k) list:
 * whole blocks, I <= K for one type.  CSR: per primitive, equal segment
 * lengths, whole partner blocks in one order for every component, block max
 * |Q| non-increasing.  vlxmi_set_list_validation(1) makes every class launcher
 * (vlxmi_launch_computeCoulombFock* / computeExchangeFock*, not the batched
 * ones) run these on its lists first and return VLXMI_ERR_LAYOUT on a
 * mismatch; off by default (each check synchronizes the stream). */
int vlxmi_validate_j_list(uint32_t la, uint32_t lb, const uint32_t* first, const uint32_t* second,
                          const double* mat_Q, const double* mat_D, uint32_t count, void* stream);
int vlxmi_validate_ik_list(uint32_t la, uint32_t lc, const uint32_t* ik_i, const uint32_t* ik_k, uint32_t count,
                           void* stream);
int vlxmi_validate_k_csr(uint32_t la, uint32_t lb, uint32_t count_a, const uint32_t* displs, const uint32_t* counts,
                         const double* Q, const uint32_t* D_inds, void* stream);
int vlxmi_set_list_validation(int on);
int vlxmi_list_validation(void);

/* Device side of the pair-list builder (vlxmi/devpairs.py drives it; VeloxChem's
 * host builds these lists for the reference kernels).  vlxmi_pair_candidates:
 * every candidate primitive-pair block (I, J) of class (la <= lb; J >= I when
 * la == lb, row-major) with I[b], J[b] and the overlap factors
 * S[b][ca][cb] = fac_a[I][ca] fac_b[J][cb] (pi/zeta)^(3/2) exp(-a_i a_j |AB|^2/zeta),
 * prim_info tables as for the kernels, fac_* per primitive component.
 * vlxmi_block_absmax: out[b] = max_t |x[b*nab + t]|.  vlxmi_kcsr_fill: exchange
 * CSR entries (Q, pair data, partner index) of ordered blocks sorted by I, block
 * max (non-increasing) and J, read from a pair set's S and Q (transposed where
 * trans[p]): entry (I, ca, rank[p], cb) at base[I] + ca*seg[I] + rank[p]*nb + cb. */
int vlxmi_pair_candidates(uint32_t la, uint32_t lb, const double* info_a, uint32_t count_a, const double* info_b,
                          uint32_t count_b, const double* fac_a, const double* fac_b, uint32_t* I, uint32_t* J,
                          double* S, void* stream);
int vlxmi_block_absmax(const double* x, uint64_t nblk, uint32_t nab, double* out, void* stream);
int vlxmi_kcsr_fill(uint32_t la, uint32_t lb, const int64_t* src, const uint8_t* trans, const int64_t* Iord,
                    const int64_t* Jord, const int64_t* rank, const int64_t* base, const int64_t* seg, uint64_t nord,
                    const double* S, const double* Q, double* Qe, double* Se, uint32_t* De, void* stream);

/* Allocation-free Coulomb launch.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same result as
 * vlxmi_launch_computeCoulombFock{CLS} with the same ket_split, for the class
 * named by cls ("ssss" .. "dddd", lower case: bra pair class then ket pair
 * class); prim_info / prim_count are indexed by angular momentum (0 s, 1 p,
 * 2 d; entries of types the class does not use are ignored), the remaining
 * arguments are the class launcher's.  With ket_split > 1 the chunks' partial
 * sums go to a static per-GPU array of vlxmi_noalloc_scratch_doubles() doubles
 * (32 MiB) instead of an allocated workspace: calls with
 * ket_split * ab_prim_pair_count_local above that return VLXMI_ERR_SIZE.  The
 * array is shared by every call on one GPU, so calls that overlap on one GPU
 * race.  The entry point calls no allocating routine. */
int vlxmi_coulomb_fock_noalloc(const char* cls, double* mat_J, const double* const* prim_info,
                               const uint32_t* prim_count, const double* cd_mat_D, const double* ab_mat_Q_local,
                               const double* cd_mat_Q, const uint32_t* ab_first_inds_local,
                               const uint32_t* ab_second_inds_local, const double* ab_pair_data_local,
                               uint32_t ab_prim_pair_count_local, const uint32_t* cd_first_inds,
                               const uint32_t* cd_second_inds, const double* cd_pair_data,
                               uint32_t cd_prim_pair_count, const double* boys_func_table,
                               const double* boys_func_ft, double eri_threshold, uint32_t ket_split, void* stream);
uint64_t vlxmi_noalloc_scratch_doubles(void);

#ifdef __cplusplus
}
#endif
#endif
