// Host-memory stand-in for the HIP runtime calls libtetemoko_hip's host code makes, for running
// that host code (box, island, tile, strip, exchange, coarse-fine and regrid list builders)
// under AddressSanitizer and UndefinedBehaviorSanitizer on a CPU.  Device allocations are host
// allocations, copies and fills are memcpy / memset, kernel launches and graph replays do
// nothing, and every call succeeds.  Nothing here computes what a kernel computes: the driver
// checks the host code's memory accesses and return codes, not results.
#include <hip/hip_runtime_api.h>

#include <cstdlib>
#include <cstring>

extern "C" {

hipError_t hipMalloc(void** p, size_t n)
{
    *p = std::aligned_alloc(256, ((n ? n : 1) + 255) / 256 * 256);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p)
{
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind)
{
    if (n) std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t)
{
    if (n) std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t)
{
    std::memset(d, v, n);
    return hipSuccess;
}
hipError_t hipMemGetInfo(size_t* free_b, size_t* total_b)
{
    *free_b = (size_t)200 << 30;
    *total_b = (size_t)288 << 30;
    return hipSuccess;
}

// handles: a distinct allocation per stream, event, graph
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned)
{
    *s = (hipStream_t)std::malloc(16);
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s)
{
    std::free(s);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamQuery(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipSuccess; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g)
{
    *g = (hipGraph_t)std::malloc(16);
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t, hipGraphNode_t*, char*, size_t)
{
    *e = (hipGraphExec_t)std::malloc(16);
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g)
{
    std::free(g);
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e)
{
    std::free(e);
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipSuccess; }

hipError_t hipEventCreate(hipEvent_t* e)
{
    *e = (hipEvent_t)std::malloc(16);
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e)
{
    std::free(e);
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t)
{
    *ms = 1.0f;
    return hipSuccess;
}

hipError_t hipGetDevice(int* d)
{
    *d = 0;
    return hipSuccess;
}
hipError_t hipGetDeviceCount(int* n)
{
    *n = 1;
    return hipSuccess;
}
hipError_t hipGetDevicePropertiesR0600(hipDeviceProp_tR0600* p, int)
{
    std::memset(p, 0, sizeof *p);
    std::strcpy(p->gcnArchName, "gfx950:sramecc+:xnack-");
    std::strcpy(p->name, "host stand-in");
    p->multiProcessorCount = 256;
    p->totalGlobalMem = (size_t)288 << 30;
    p->warpSize = 64;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int)
{
    *v = a == hipDeviceAttributeMultiprocessorCount ? 256 : 64;
    return hipSuccess;
}
hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return hipSuccess; }

static char g_symbols[64 << 20];  // device symbols (the allocation-free entry point's scratch)
hipError_t hipGetSymbolAddress(void** p, const void*)
{
    *p = g_symbols;
    return hipSuccess;
}

hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "host stand-in"; }

hipError_t hipLaunchKernel(const void*, dim3, dim3, void**, size_t, hipStream_t) { return hipSuccess; }

// the compiler's launch and registration hooks
static dim3 g_grid, g_block;
static size_t g_shmem;
static hipStream_t g_stream;
hipError_t __hipPushCallConfiguration(dim3 grid, dim3 block, size_t shmem, hipStream_t stream)
{
    g_grid = grid;
    g_block = block;
    g_shmem = shmem;
    g_stream = stream;
    return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3* grid, dim3* block, size_t* shmem, hipStream_t* stream)
{
    *grid = g_grid;
    *block = g_block;
    *shmem = g_shmem;
    *stream = g_stream;
    return hipSuccess;
}
static void* g_module[4];
void** __hipRegisterFatBinary(const void*) { return g_module; }
void __hipRegisterFunction(void**, const void*, char*, const char*, unsigned, void*, void*, dim3*, dim3*, int*) {}
void __hipRegisterVar(void**, void*, char*, char*, int, size_t, int, int) {}
void __hipUnregisterFatBinary(void**) {}

}  // extern "C"
