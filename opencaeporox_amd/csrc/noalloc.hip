// noalloc.hip — ocpx_phase_equilibrium_noalloc: the batched phase
// equilibrium of ocpx_phase_equilibrium for callers that cannot allocate
// device memory.  The caller passes the inputs and outputs in device memory;
// the kernel's scratch -- the work-queue counter and the device copy of the
// fluid constants -- lives in one file-scope device array of fixed size, so
// the call makes no hipMalloc, hipFree or other allocating call.
//
// Each device holds its own copy of the array for the life of the process;
// its address on a device is taken once, with that device current, through
// hipGetSymbolAddress, and used from then on in place of the array's name.
// Every call on a device shares that device's copy: calls that overlap on
// one device race (the counter and the constants would be overwritten).
#include <algorithm>
#include <mutex>
#include <string>

#include "ocpx_dev.hpp"
#include "ocpx_kernels.hpp"

namespace ocpx {
void set_call_error(const std::string& msg);
void clear_call_error();
}  // namespace ocpx

using namespace ocpx;

namespace {

constexpr size_t SCRATCH_BYTES = 8192;     // queue counter (first 256 bytes), then the fluid
constexpr size_t FLUID_OFF = 256;
static_assert(FLUID_OFF + sizeof(DevFluid) <= SCRATCH_BYTES, "the fluid constants must fit the scratch array");

__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
unsigned char* g_ptr[MAX_DEVICES] = {};   // scratch's address on each device
int g_grid[MAX_DEVICES][OCPX_MAX_NC + 1] = {};

// bytes of a host object carried to the device as a kernel argument (no host
// buffer has to outlive the call)
struct Chunk {
    uint32_t off, len;
    unsigned char b[1024];
};

__global__ void __launch_bounds__(256) put_bytes(unsigned char* base, Chunk c)
{
    for (uint32_t k = threadIdx.x; k < c.len; k += blockDim.x) base[c.off + k] = c.b[k];
}

__global__ void __launch_bounds__(64) zero_word(unsigned* p)
{
    if (threadIdx.x == 0) *p = 0u;
}

int fail(int code, const std::string& msg)
{
    set_call_error("ocpx_phase_equilibrium_noalloc: " + msg);
    return code;
}

// the scratch array's address on the current device (looked up once per device)
int scratch_of_device(unsigned char** out, int* dev_out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return fail(OCPX_E_HIP, "no current device");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(OCPX_E_HIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_ptr[dev] = (unsigned char*)p;
    }
    *out = g_ptr[dev];
    *dev_out = dev;
    return OCPX_OK;
}

}  // namespace

extern "C" int ocpx_phase_equilibrium_noalloc(const ocpx_fluid* fluid, int64_t n, const double* P, const double* T,
                                              const double* z, int32_t* NP, double* nu, double* x, int32_t* ftype,
                                              void* hip_stream)
{
    clear_call_error();
    if (!fluid) return fail(OCPX_E_ARG, "null fluid");
    if (fluid->nc < 2 || fluid->nc > 12) return fail(OCPX_E_ARG, "Nc must be 2..12");
    if (n < 0 || n > INT32_MAX) return fail(OCPX_E_ARG, "n must be 0..2^31-1 cells");
    if (n == 0) return OCPX_OK;
    if (!P || !T || !z || !NP || !nu || !x || !ftype) return fail(OCPX_E_ARG, "null array");
    unsigned char* base = nullptr;
    int dev = 0;
    int rc = scratch_of_device(&base, &dev);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)hip_stream;
    // the fluid constants, SkipPSA off: every cell runs the strict test from
    // one phase, as PhaseEquilibrium(P, T, z) with ftype 0
    DevFluid h;
    build_dev_fluid(fluid, h);
    h.skip = 0;
    const unsigned char* hb = reinterpret_cast<const unsigned char*>(&h);
    for (uint32_t off = 0; off < sizeof(DevFluid); off += sizeof(Chunk::b)) {
        Chunk c;
        c.off = (uint32_t)FLUID_OFF + off;
        c.len = (uint32_t)std::min(sizeof(Chunk::b), sizeof(DevFluid) - off);
        for (uint32_t k = 0; k < c.len; k++) c.b[k] = hb[off + k];
        hipLaunchKernelGGL(put_bytes, dim3(1), dim3(256), 0, st, base, c);
    }
    unsigned* queue = reinterpret_cast<unsigned*>(base);
    hipLaunchKernelGGL(zero_word, dim3(1), dim3(64), 0, st, queue);
    if (hipGetLastError() != hipSuccess) return fail(OCPX_E_HIP, "scratch set-up launch failed");
    PeIO io{};
    io.nb = n;
    io.P = P;
    io.T = T;
    io.Ni = z;                     // normalised again by the kernel (InitNtZ)
    io.pe_NP = NP;
    io.pe_nu = nu;
    io.pe_x = x;
    io.ftype_out = ftype;
    io.queue = queue;
    int grid;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_grid[dev][fluid->nc]) {
            int ncu = 0, per = 0;
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
            per = pe_batched_blocks_per_cu(fluid->nc);
            g_grid[dev][fluid->nc] = (ncu > 0 ? ncu : 256) * (per > 0 ? per : 2);
        }
        grid = g_grid[dev][fluid->nc];
    }
    const hipError_t e = run_pe(fluid->nc, reinterpret_cast<const DevFluid*>(base + FLUID_OFF), io,
                                OCPX_FLASH_BATCHED, grid, st);
    if (e != hipSuccess) return fail(OCPX_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return OCPX_OK;
}
