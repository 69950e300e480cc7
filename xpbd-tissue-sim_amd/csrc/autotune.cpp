// Autotuner: times the kernel variants and launch shapes that share the
// handle's device data layout on the handle's own problem, keeps the fastest,
// and caches the choice per problem shape in a small text file.
//
// Gauss-Seidel handles built with clusters can switch between the one-thread
// cluster kernel and the quad kernel at 64, 128 or 256 threads per workgroup
// without rebuilding anything (same clusters, colours and order, so the same
// results up to FP32 summation order within an element).  Patch Jacobi handles
// time the fused one-launch iteration against the patch + shared-vertex +
// contact launches (the same sums; with contacts the contact corrections are
// added in another order).  Cluster and patch kernels are also timed with and
// without the XCD-contiguous workgroup mapping (which does not change results).
// Each candidate's iterations are captured into a HIP graph and its replays are
// timed.  Other handles have a single candidate; the call then only records it.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "capi_internal.h"
#include "kernels.h"

using namespace xpbd;

namespace {

struct Candidate {
    int variant;
    int block;
    int xcd;
};

std::string shape_key(const xpbd_handle* h)
{
    std::ostringstream k;
    k << (h->params.solver == XPBD_SOLVER_GS_COLORED ? "gs" : "jacobi") << ":v" << h->nv << ":t" << h->nt << ":c"
      << h->ncolors << ":k" << h->nclusters << ":i" << h->ninst << ":x";
    return k.str();
}

bool read_cache(const char* path, const std::string& key, Candidate* out)
{
    if (!path) return false;
    std::ifstream f(path);
    std::string line;
    bool found = false;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        std::string k;
        int v = 0, b = 0, x = 0;
        if (is >> k >> v >> b >> x && k == key) {
            *out = { v, b, x };
            found = true;  // the last entry for a key wins
        }
    }
    return found;
}

void write_cache(const char* path, const std::string& key, const Candidate& c, double ms_per_iter)
{
    if (!path) return;
    std::ofstream f(path, std::ios::app);
    f << key << " " << c.variant << " " << c.block << " " << c.xcd << " " << ms_per_iter << "\n";
}

bool valid(const xpbd_handle* h, const Candidate& c)
{
    if (c.variant == XPBD_VARIANT_CLUSTER) return h->nclusters > 0;
    if (c.variant == XPBD_VARIANT_CLUSTER_QUAD) return h->nclusters > 0 && quad_block_supported(c.block);
    if (c.variant == XPBD_VARIANT_PATCH) return h->variant == XPBD_VARIANT_PATCH && (c.block == 0 || h->fused_ok);
    return c.variant == h->variant;
}

void apply(xpbd_handle* h, const Candidate& c)
{
    if (h->variant != c.variant || (c.variant == XPBD_VARIANT_CLUSTER_QUAD && h->quad_block != c.block) ||
        (c.variant == XPBD_VARIANT_PATCH && h->patch_fused != (c.block == 1)) || h->xcd_map != (c.xcd != 0)) {
        h->variant = c.variant;
        if (c.variant == XPBD_VARIANT_CLUSTER_QUAD) h->quad_block = c.block;
        if (c.variant == XPBD_VARIANT_PATCH) h->patch_fused = c.block == 1;
        h->xcd_map = c.xcd != 0;
        invalidate_graphs(h);
    }
    char buf[64];
    std::snprintf(buf, sizeof buf, "variant %d block %d xcd %d", c.variant,
                  c.variant == XPBD_VARIANT_CLUSTER_QUAD || c.variant == XPBD_VARIANT_PATCH ? c.block : 0, c.xcd);
    h->tuned = buf;
}

}  // namespace

extern "C" int xpbd_autotune(xpbd_batch* h, const char* cache_path)
{
    if (!h) return XPBD_E_INVALID_ARG;
    const std::string key = shape_key(h);
    Candidate c{};
    if (read_cache(cache_path, key, &c) && valid(h, c)) {
        apply(h, c);
        return XPBD_OK;
    }
    std::vector<Candidate> cand;
    if (h->params.solver == XPBD_SOLVER_GS_COLORED && h->nclusters > 0) {
        for (int x = 1; x >= 0; x--)
            for (const Candidate& b : { Candidate{ XPBD_VARIANT_CLUSTER, 0, x }, Candidate{ XPBD_VARIANT_CLUSTER_QUAD, 64, x },
                                        Candidate{ XPBD_VARIANT_CLUSTER_QUAD, 128, x },
                                        Candidate{ XPBD_VARIANT_CLUSTER_QUAD, 256, x },
                                        Candidate{ XPBD_VARIANT_CLUSTER_QUAD, -64, x },    // LDS-lean forms
                                        Candidate{ XPBD_VARIANT_CLUSTER_QUAD, -128, x },
                                        Candidate{ XPBD_VARIANT_CLUSTER_QUAD, -256, x } }) {
                cand.push_back(b);
                // the same quad shape with the split delta layout (not with attachments)
                if (b.variant == XPBD_VARIANT_CLUSTER_QUAD && h->natt == 0)
                    cand.push_back({ b.variant, b.block < 0 ? b.block - kQuadSplit : b.block + kQuadSplit, x });
            }
    } else if (h->variant == XPBD_VARIANT_PATCH) {
        // block 1: the fused one-launch iteration (when the layout allows it)
        for (int f = h->fused_ok ? 1 : 0; f >= 0; f--)
            for (int x = 1; x >= 0; x--) cand.push_back({ h->variant, f, x });
    } else {
        cand = { { h->variant, h->quad_block, h->xcd_map ? 1 : 0 } };
    }
    if (cand.size() == 1) {
        apply(h, cand[0]);
        write_cache(cache_path, key, cand[0], 0.0);
        return XPBD_OK;
    }
    // snapshot what the trial iterations change: vertex records and multipliers
    hipStream_t s = h->stream;
    const size_t pos_bytes = sizeof(float4) * 2 * (size_t)h->nv, lam_bytes = sizeof(float2) * (size_t)h->nt;
    void *save_pos = nullptr, *save_lam = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&save_pos, pos_bytes);
    if (e == hipSuccess) e = hipMalloc(&save_lam, lam_bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemcpyAsync(save_pos, h->d_pos, pos_bytes, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(save_lam, h->d_lam, lam_bytes, hipMemcpyDeviceToDevice, s);
    int rc = e == hipSuccess ? XPBD_OK : hip_err(h, e, "autotune set-up");
    const int saved_variant = h->variant, saved_block = h->quad_block;
    const bool saved_xcd = h->xcd_map, saved_fused = h->patch_fused;
    double best_ms = 1e300;
    Candidate best = cand[0];
    const int iters = h->params.iters > 0 ? h->params.iters : 1;
    for (const Candidate& cd : cand) {
        if (rc) break;
        h->variant = cd.variant;
        if (cd.variant == XPBD_VARIANT_CLUSTER_QUAD) h->quad_block = cd.block;
        if (cd.variant == XPBD_VARIANT_PATCH) h->patch_fused = cd.block == 1;
        h->xcd_map = cd.xcd != 0;
        double ms_best = 1e300;
        // the candidate's iterations are captured into a graph and the replays are
        // timed, as the solver replays its steps: launches issued one by one from
        // the host would time the host where its per-launch cost exceeds a kernel's
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        e = hipStreamSynchronize(s);
        if (e == hipSuccess) e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            rc = hip_err(h, e, "autotune capture");
            break;
        }
        rc = enqueue_iterations(h, iters, false);
        e = hipStreamEndCapture(s, &g);
        if (!rc && e != hipSuccess) rc = hip_err(h, e, "autotune capture");
        if (!rc) {
            e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            if (e != hipSuccess) rc = hip_err(h, e, "autotune graph");
        }
        if (g) (void)hipGraphDestroy(g);
        for (int rep = 0; rep < 4 && !rc; rep++) {  // first repetition warms up
            (void)hipEventRecord(e0, s);
            e = hipGraphLaunch(ge, s);
            if (e != hipSuccess) rc = hip_err(h, e, "autotune replay");
            (void)hipEventRecord(e1, s);
            if (!rc) {
                e = hipEventSynchronize(e1);
                float ms = 0.f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (e != hipSuccess) rc = hip_err(h, e, "autotune timing");
                else if (rep > 0 && ms < ms_best) ms_best = ms;
            }
            // every trial starts from the same state
            if (!rc) e = hipMemcpyAsync(h->d_pos, save_pos, pos_bytes, hipMemcpyDeviceToDevice, s);
            if (!rc && e == hipSuccess) e = hipMemcpyAsync(h->d_lam, save_lam, lam_bytes, hipMemcpyDeviceToDevice, s);
            if (!rc && e != hipSuccess) rc = hip_err(h, e, "autotune restore");
        }
        if (ge) {
            (void)hipStreamSynchronize(s);
            (void)hipGraphExecDestroy(ge);
        }
        if (!rc && ms_best < best_ms) {
            best_ms = ms_best;
            best = cd;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    if (save_pos) (void)hipFree(save_pos);
    if (save_lam) (void)hipFree(save_lam);
    h->variant = saved_variant;
    h->quad_block = saved_block;
    h->xcd_map = saved_xcd;
    h->patch_fused = saved_fused;
    if (rc) return rc;
    apply(h, best);
    write_cache(cache_path, key, best, best_ms / iters);
    return XPBD_OK;
}

extern "C" int xpbd_get_variant(xpbd_handle* h, int32_t* variant, int32_t* block, int32_t* xcd)
{
    if (!h || !variant) return XPBD_E_INVALID_ARG;
    *variant = h->variant;
    if (block)
        *block = h->variant == XPBD_VARIANT_CLUSTER_QUAD ? h->quad_block
               : h->variant == XPBD_VARIANT_PATCH    ? (h->patch_fused && h->fused_ok ? 1 : 0)
                                                     : 0;
    if (xcd) *xcd = h->xcd_map ? 1 : 0;
    return XPBD_OK;
}
