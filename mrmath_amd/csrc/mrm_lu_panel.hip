// mrm_lu_panel.hip -- the launcher of the LU panel kernel
// (mrm_lu_panel_kernel.h; one translation unit per rows-per-thread variant).

#include "mrm_lu_panel_kernel.h"

namespace {
int panel_grid(int64_t n, int b) { return int((n + b - 1) / b); }
}  // namespace

namespace mrm {

template <typename T>
int lu_panel(T* P, int64_t lda, int m, int w, int piv_base, int* indx, int* perm, LuWork& ws, hipStream_t s) {
    constexpr int W = LU_W;
    // Every column reads every workgroup's record, so rows are packed up to 4
    // per thread before more workgroups are added.
    int rpt = 1;
    int G = panel_grid(m, PANEL_THREADS);
    while (G > LU_PANEL_WG_TARGET && rpt < 4) {
        rpt *= 2;
        G = panel_grid(m, rpt * PANEL_THREADS);
    }
    if (G > LU_MAX_PANEL_WG) {
        mrm_set_error(MRM_BAD_DIMS, "lu: matrix too tall for the panel kernel (%d rows)", m);
        return MRM_BAD_DIMS;
    }
    uint64_t* xch = static_cast<uint64_t*>(ws.xch);
    // flags are unique within a factorisation (the exchange area is zeroed at
    // its start): column j of this panel is flagged piv_base + j + 1
    uint32_t flag_base = uint32_t(piv_base) + 1u;
    uint64_t timeout = ws.coop ? LU_PANEL_TIMEOUT_TICKS : ws.plain_timeout;
    void* args[] = {&P, &lda, &m, &w, &piv_base, &indx, &xch, &flag_base, &ws.status, &timeout};
    hipError_t e;
    const int ri = rpt == 1 ? 0 : rpt == 2 ? 1 : 2;
    static_assert(sizeof(T) == sizeof(double) && W == LU_W, "the variants are compiled for FP64 panels");
    const void* fn = ri == 0 ? lu_panel_fn<1>() : ri == 1 ? lu_panel_fn<2>() : lu_panel_fn<4>();
    // The workgroups wait for each other, so all G must be resident at once.
    // A plain launch keeps them so whenever the device can hold G of them and
    // no other work holds its CUs: the stream's earlier kernels have
    // finished, and work on other streams does not depend on this kernel.  It
    // saves the ~20 us of queue time a cooperative launch costs per panel.
    // If other work did hold CUs, a wait passes its deadline (never a hang),
    // the factorisation reports MRM_LU_TIMEOUT and lu_factor_sync relaunches
    // it with ws.coop set: every panel then takes the cooperative launch,
    // which guarantees co-residency.  Larger grids are always cooperative.
    // (Callers hold the host mutex.)
    // On the lookahead's panel stream only ws.panel_cus CUs are usable, and
    // there the launch is always plain.
    constexpr int MAX_DEV = 64;
    static int per_cu_of[MAX_DEV][3], cus_of[MAX_DEV];
    static bool known[MAX_DEV][3];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEV) dev = -1;
    int cap = 0;
    if (dev >= 0) {
        if (!known[dev][ri]) {
            int per_cu = 0, cus = 0;
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, PANEL_THREADS, 0) != hipSuccess)
                per_cu = cus = 0;
            per_cu_of[dev][ri] = per_cu;
            cus_of[dev] = cus;
            known[dev][ri] = true;
        }
        cap = per_cu_of[dev][ri] * (ws.panel_cus ? ws.panel_cus : cus_of[dev]);
    }
    if (ws.panel_cus && (ws.coop || G > cap)) {
        mrm_set_error(MRM_BAD_DIMS, "lu panel: %d workgroups do not fit the lookahead's %d CUs", G, ws.panel_cus);
        return MRM_BAD_DIMS;
    }
    if (!ws.coop && G <= cap)
        e = hipLaunchKernel(fn, dim3(G), dim3(PANEL_THREADS), args, 0, s);
    else
        e = hipLaunchCooperativeKernel(fn, dim3(G), dim3(PANEL_THREADS), args, 0, s);
    if (e != hipSuccess) {
        mrm_set_error(MRM_HIP_ERROR, "lu panel launch (G=%d): %s", G, hipGetErrorString(e));
        return MRM_HIP_ERROR;
    }
    return MRM_OK;
}

template int lu_panel<double>(double*, int64_t, int, int, int, int*, int*, LuWork&, hipStream_t);

size_t lu_xch_bytes() {
    // two parities of G flagged records plus row j, 8-byte words, G <= LU_MAX_PANEL_WG
    return 2 * Rec<double, LU_W>::parity_units(LU_MAX_PANEL_WG) * sizeof(uint64_t);
}

}  // namespace mrm
