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
