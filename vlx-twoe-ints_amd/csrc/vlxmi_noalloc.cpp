// Coulomb launches that allocate nothing: the ket-split partial sums live in a
// static device array instead of the per-stream workspace the public launchers
// allocate.  For callers that cannot call hipMalloc (HIP graph capture, tight
// serving loops, harnesses that forbid allocation).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>

#include "vlxmi_device.hpp"

// Partial-sum scratch: VLXMI_NOALLOC_SCRATCH_DOUBLES doubles (32 MiB) per GPU,
// held by every GPU the library runs on until the process ends.  Every call on
// one GPU shares it, so calls that overlap on one GPU (different streams, or
// work still running from an earlier call on another stream) race.  Each call
// writes every slot it later reads (every chunk stores all of its partial
// sums), so nothing needs clearing between calls.
#define VLXMI_NOALLOC_SCRATCH_DOUBLES (4u << 20)
static __device__ __align__(256) double scratch[VLXMI_NOALLOC_SCRATCH_DOUBLES];

extern "C" int vlxmi_jpart_dispatch(const char* cls, double* mat_J, const double* const* prim_info,
                                    const uint32_t* prim_count, const double* cd_mat_D, const double* ab_mat_Q_local,
                                    const double* cd_mat_Q, const uint32_t* ab_first_inds_local,
                                    const uint32_t* ab_second_inds_local, const double* ab_pair_data_local,
                                    uint32_t ab_prim_pair_count_local, const uint32_t* cd_first_inds,
                                    const uint32_t* cd_second_inds, const double* cd_pair_data,
                                    uint32_t cd_prim_pair_count, const double* boys_func_table,
                                    const double* boys_func_ft, double eri_threshold, uint32_t ket_split,
                                    double* part, void* stream);

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mutex;
double* g_scratch[kMaxDevices] = {};

// The device address of this GPU's copy of ``scratch``, looked up once per
// device (with that device current).  Host code uses only this address; the
// array's name is a host-side stand-in.
int scratch_address(double** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kMaxDevices) return VLXMI_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        double* p = nullptr;
        e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return VLXMI_OK;
}

int ncart_of(char x) { return x == 's' ? 1 : x == 'p' ? 3 : x == 'd' ? 6 : 0; }
}  // namespace

extern "C" uint64_t vlxmi_noalloc_scratch_doubles(void) { return VLXMI_NOALLOC_SCRATCH_DOUBLES; }

extern "C" int vlxmi_coulomb_fock_noalloc(const char* cls, double* mat_J, const double* const* prim_info,
                                          const uint32_t* prim_count, const double* cd_mat_D,
                                          const double* ab_mat_Q_local, const double* cd_mat_Q,
                                          const uint32_t* ab_first_inds_local, const uint32_t* ab_second_inds_local,
                                          const double* ab_pair_data_local, uint32_t ab_prim_pair_count_local,
                                          const uint32_t* cd_first_inds, const uint32_t* cd_second_inds,
                                          const double* cd_pair_data, uint32_t cd_prim_pair_count,
                                          const double* boys_func_table, const double* boys_func_ft,
                                          double eri_threshold, uint32_t ket_split, void* stream)
{
    if (!cls || std::strlen(cls) != 4 || !prim_info || !prim_count) return VLXMI_ERR_ARG;
    const int nc = ncart_of(cls[2]), nd = ncart_of(cls[3]);
    if (!ncart_of(cls[0]) || !ncart_of(cls[1]) || !nc || !nd) return VLXMI_ERR_ARG;
    // the split the launch will use (as the class launcher clamps it)
    const uint32_t nket = cd_prim_pair_count / (uint32_t)(nc * nd);
    if (ket_split == 0) ket_split = 1;
    if (nket > 0 && ket_split > nket) ket_split = nket;
    double* part = nullptr;
    if (ket_split > 1) {
        if ((uint64_t)ket_split * ab_prim_pair_count_local > VLXMI_NOALLOC_SCRATCH_DOUBLES) return VLXMI_ERR_SIZE;
        const int rc = scratch_address(&part);
        if (rc != 0) return rc;
    }
    return vlxmi_jpart_dispatch(cls, mat_J, prim_info, prim_count, cd_mat_D, ab_mat_Q_local, cd_mat_Q,
                                ab_first_inds_local, ab_second_inds_local, ab_pair_data_local,
                                ab_prim_pair_count_local, cd_first_inds, cd_second_inds, cd_pair_data,
                                cd_prim_pair_count, boys_func_table, boys_func_ft, eri_threshold, ket_split, part,
                                stream);
}
