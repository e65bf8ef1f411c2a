// mf_noalloc.hip -- mf_solve_noalloc: the plain solve of mf_solve without any
// device allocation, for callers that cannot allocate (code under HIP graph
// capture elsewhere in the process, tight serving loops, harnesses that
// forbid hipMalloc).
//
// The query's slab (residual, excess, heights, frontiers and lists, marks)
// is carved from `scratch`, one statically sized array per device, instead
// of the engine's slab arena.  The query runs on its full graph: pruning
// would build and cache induced subgraphs (device allocations), and captured
// fused cycles would be instantiated (another), so both are off on this
// path; the flow is the same, certified as in mf_solve.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/mscmf.h"
#include "mf_engine.h"
#include "mf_handles.h"

namespace {

constexpr size_t kScratchBytes = (size_t)MF_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

// Each device the library runs on holds its own copy until the process ends.
// Host code never uses this name as an address: it uses the per-device
// pointer from hipGetSymbolAddress below.
static __device__ __align__(256) char scratch[kScratchBytes];

struct DevScratch {
    std::once_flag once;
    hipError_t rc = hipSuccess;
    char* ptr = nullptr;
    std::mutex use;  // calls on one device share its array: one at a time
};
DevScratch g_dev[kMaxDevices];

// The device address of `scratch` on the current device, looked up once per
// device with that device current.
int scratch_of(int dev, char** out) {
    if (dev < 0 || dev >= kMaxDevices) {
        mf::set_error("device index out of range for the no-allocation scratch");
        return MF_EINVAL;
    }
    DevScratch& d = g_dev[dev];
    std::call_once(d.once, [&] {
        void* p = nullptr;
        d.rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        d.ptr = (char*)p;
    });
    if (d.rc != hipSuccess || !d.ptr) {
        mf::set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(d.rc));
        return MF_EHIP;
    }
    *out = d.ptr;
    return MF_OK;
}

}  // namespace

extern "C" int mf_solve_noalloc(mf_engine* e, int gid, int s, int t, mf_result* r) {
    if (!e) { mf::set_error("null engine"); return MF_EINVAL; }
    if (gid < 0 || gid >= e->eng.ngraphs()) { mf::set_error("bad graph index"); return MF_EINVAL; }
    const int n = e->eng.graph(gid)->g.n;
    if (s < 0 || t < 0 || s >= n || t >= n) { mf::set_error("vertex out of range"); return MF_EINVAL; }
    if (s == t) { mf::set_error("s == t"); return MF_EINVAL; }
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) { mf::set_error("no HIP device"); return MF_ENODEV; }
    const int dev = e->eng.device();
    if (cur != dev && hipSetDevice(dev) != hipSuccess) { mf::set_error("cannot select the engine's device"); return MF_EHIP; }
    char* base = nullptr;
    int rc = scratch_of(dev, &base);
    if (rc == MF_OK) {
        std::lock_guard<std::mutex> lk(g_dev[dev].use);
        mf::Query q;
        q.graph = gid;
        q.s = s;
        q.t = t;
        q.tag = -1;
        std::vector<mf::Result> out;
        out.reserve(1);
        rc = e->eng.solve_in(q, base, kScratchBytes, out);
        if (rc == MF_OK && out.size() != 1) { mf::set_error("solve produced no result"); rc = MF_EHIP; }
        if (rc == MF_OK) {
            if (r) mf::result_to_c(out[0], r);
            rc = out[0].st.status;
        }
    }
    if (cur != dev) (void)hipSetDevice(cur);
    return rc;
}
