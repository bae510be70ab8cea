// cc_noalloc.hip -- ccgpu_energy_and_all_force_noalloc: the energy and every force of one configuration with all
// the device memory of the evaluation (particle arrays, constraint tables, descriptor, work list, partial sums and
// results) carved from a fixed per-device scratch array of this library, for callers that cannot allocate device
// memory.  The kernels, tiles and summation order are those of a bound potential, so the results carry the same
// bits as ccgpu_potential_energy_and_all_force on the same configuration and constraints.
#include "cc_abi.h"

#include <cmath>
#include <mutex>

using namespace ccgpu;

namespace {
constexpr size_t kScratchBytes = (size_t)64 << 20;  // the stated maximum of one evaluation's device memory
constexpr int kMaxDevices = 64;
}  // namespace

// One copy per GPU the library runs on, held until the process ends.  On the host the name only stands for the
// symbol: host code passes the device address that hipGetSymbolAddress returns for the current device, never the
// name itself.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
struct DeviceScratch {
    std::mutex mu;        // the calls on one device share its array: they run one at a time
    bool ready = false;
    void *ptr = nullptr;  // device address of this device's copy of the array
};
DeviceScratch g_scratch[kMaxDevices];

// set-up, on the first call on a device (that device current); S.mu is held
void *ScratchAddress(DeviceScratch &S)
{
    if (!S.ready) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            (void)hipGetLastError();
            throw Error(CCGPU_ERR_DEVICE,
                        std::string("hipGetSymbolAddress of the scratch array failed: ") + hipGetErrorString(e));
        }
        S.ptr = p;
        S.ready = true;
    }
    return S.ptr;
}
}  // namespace

extern "C" int ccgpu_energy_and_all_force_noalloc(int dim, const double *basis, int64_t N, const double *relative,
                                                  int64_t M, const double *k, const double *V, double *E, double *F)
{
    return Guard([&] {
        Require(dim >= 1 && dim <= kMaxDim, CCGPU_ERR_ARG, "dimension must be 1..4");
        Require(basis && E && N >= 0 && M >= 0 && (N == 0 || (relative && F)) && (M == 0 || (k && V)), CCGPU_ERR_ARG,
                "bad argument");
        Require(N < (int64_t)1 << 30, CCGPU_ERR_ARG, "too many particles");
        for (int i = 0; i < dim * dim; i++) Require(std::isfinite(basis[i]), CCGPU_ERR_ARG, "non-finite basis");
        for (int64_t i = 0; i < N * dim; i++)
            if (!std::isfinite(relative[i])) throw Error(CCGPU_ERR_NAN, "non-finite particle coordinate");
        RequireDevice();
        int dev = 0;
        CC_HIP(hipGetDevice(&dev));
        Require(dev >= 0 && dev < kMaxDevices, CCGPU_ERR_DEVICE, "device index beyond the scratch table");
        DeviceScratch &S = g_scratch[dev];
        std::lock_guard<std::mutex> lk(S.mu);
        Arena arena;
        arena.base = static_cast<char *>(ScratchAddress(S));
        arena.cap = kScratchBytes;
        ArenaScope scope(arena);  // every DevArray reserved below is carved from the array, none from hipMalloc
        System sys(dim);
        sys.alone = true;  // the tiles of a bound potential
        Batch batch(nullptr);  // the null stream: no stream is created
        batch.systems.push_back(&sys);
        sys.SetCell(basis);
        sys.SetConstraints(M, k, M > 0 ? V : nullptr);
        sys.N = N;
        batch.Build();  // an over-size evaluation throws here, before any kernel runs
        if (N > 0)
            CC_HIP(hipMemcpyAsync(sys.x.get(), relative, (size_t)(N * dim) * sizeof(double), hipMemcpyHostToDevice,
                                  batch.stream));
        batch.Enqueue(true, true);
        double e = 0;
        CC_HIP(hipMemcpyAsync(&e, sys.energy.get(), sizeof(double), hipMemcpyDeviceToHost, batch.stream));
        if (N > 0)
            CC_HIP(hipMemcpyAsync(F, sys.F.get(), (size_t)(N * dim) * sizeof(double), hipMemcpyDeviceToHost,
                                  batch.stream));
        batch.Sync();
        *E = e;
    });
}
