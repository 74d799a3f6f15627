// Edge material indices from cell materials on the device: the classification
// half of build_electric_components / build_magnetic_components
// (gprMax/yee_cell_build_ext.pyx:110-257).
//
// An edge (E) or face (H) takes the material of the cells around it: four
// cells for E, two for H, in the reference's numID1.. order.  Edges of a
// rigid cell keep the index the geometry gave them; edges whose cells agree
// take that material; the others need an averaged material, which only the
// host can find or create (it walks the material list).  The device collects
// the distinct sets of cell materials of those edges in a hash table, with the
// first edge (in the reference's loop order) of each; the host resolves each
// set once, in first-edge order, and the device writes the result to every
// edge of the set.
//
// Keys are the cell materials sorted and packed 16 bits each (indices below
// 65535); slot keys start at KEY_EMPTY.  A full table makes gpx_edges_classify
// fail with GPX_ELIMIT after the kernel has finished.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "gpx.h"

namespace {

constexpr unsigned long long KEY_EMPTY = ~0ull;

struct EdgeArgs {
    int nx, ny, nz;
    int comp;                       // 0..5: Ex, Ey, Ez, Hx, Hy, Hz
    const uint32_t *solid;          // [nx][ny][nz]
    const uint8_t *rigid;           // [nx][ny][nz], E or H flags
    uint32_t *ID;                   // [6][nx+1][ny+1][nz+1]
    unsigned long long *keys;       // [cap]
    unsigned long long *first;      // [cap]: smallest edge index of the set
    unsigned int cap;               // power of two
    unsigned int *overflow;
    const uint32_t *resolved;       // [cap] (scatter pass)
};

// The cells around edge (i, j, k) of component `comp`, in numID1.. order
// (yee_cell_build_ext.pyx:127-130, 149-152, 171-174, 205-206, 225-226, 245-246);
// returns how many (4 or 2), or 0 when the edge is outside the loop ranges.
__device__ __forceinline__ int edge_cells(const EdgeArgs &a, int i, int j, int k, int c[4][3])
{
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    switch (a.comp) {
    case 0:
        if (i >= nx || j < 1 || j >= ny || k < 1 || k >= nz) return 0;
        c[0][0] = i; c[0][1] = j;     c[0][2] = k;
        c[1][0] = i; c[1][1] = j - 1; c[1][2] = k;
        c[2][0] = i; c[2][1] = j - 1; c[2][2] = k - 1;
        c[3][0] = i; c[3][1] = j;     c[3][2] = k - 1;
        return 4;
    case 1:
        if (i < 1 || i >= nx || j >= ny || k < 1 || k >= nz) return 0;
        c[0][0] = i;     c[0][1] = j; c[0][2] = k;
        c[1][0] = i - 1; c[1][1] = j; c[1][2] = k;
        c[2][0] = i - 1; c[2][1] = j; c[2][2] = k - 1;
        c[3][0] = i;     c[3][1] = j; c[3][2] = k - 1;
        return 4;
    case 2:
        if (i < 1 || i >= nx || j < 1 || j >= ny || k >= nz) return 0;
        c[0][0] = i;     c[0][1] = j;     c[0][2] = k;
        c[1][0] = i - 1; c[1][1] = j;     c[1][2] = k;
        c[2][0] = i - 1; c[2][1] = j - 1; c[2][2] = k;
        c[3][0] = i;     c[3][1] = j - 1; c[3][2] = k;
        return 4;
    case 3:
        if (i < 1 || i >= nx || j >= ny || k >= nz) return 0;
        c[0][0] = i; c[0][1] = j; c[0][2] = k;
        c[1][0] = i - 1; c[1][1] = j; c[1][2] = k;
        return 2;
    case 4:
        if (i >= nx || j < 1 || j >= ny || k >= nz) return 0;
        c[0][0] = i; c[0][1] = j; c[0][2] = k;
        c[1][0] = i; c[1][1] = j - 1; c[1][2] = k;
        return 2;
    default:
        if (i >= nx || j >= ny || k < 1 || k >= nz) return 0;
        c[0][0] = i; c[0][1] = j; c[0][2] = k;
        c[1][0] = i; c[1][1] = j; c[1][2] = k - 1;
        return 2;
    }
}

// Packs the sorted materials of a mixed edge, or returns KEY_EMPTY for an
// edge that is skipped (outside the ranges, rigid) or uniform (then *uniform
// holds its material).
__device__ __forceinline__ unsigned long long edge_key(const EdgeArgs &a, int i, int j, int k, uint32_t *uniform)
{
    int c[4][3];
    const int n = edge_cells(a, i, j, k, c);
    *uniform = 0xffffffffu;
    if (!n) return KEY_EMPTY;
    uint32_t m[4];
    bool rigid = false;
    for (int q = 0; q < n; ++q) {
        const size_t o = ((size_t)c[q][0] * a.ny + c[q][1]) * a.nz + c[q][2];
        m[q] = a.solid[o];
        rigid = rigid || a.rigid[o];
    }
    if (rigid) return KEY_EMPTY;
    bool same = true;
    for (int q = 1; q < n; ++q) same = same && m[q] == m[0];
    if (same) {
        *uniform = m[0];
        return KEY_EMPTY;
    }
    for (int p = 1; p < n; ++p)  // insertion sort of at most four values
        for (int q = p; q > 0 && m[q - 1] > m[q]; --q) {
            const uint32_t t = m[q];
            m[q] = m[q - 1];
            m[q - 1] = t;
        }
    unsigned long long key = 0;
    for (int q = 0; q < n; ++q) key = (key << 16) | m[q];
    return key;
}

__device__ __forceinline__ unsigned int key_hash(unsigned long long key, unsigned int cap)
{
    key ^= key >> 33;
    key *= 0xff51afd7ed558ccdull;
    key ^= key >> 33;
    return (unsigned int)key & (cap - 1);
}

// Pass 1: uniform edges get their material; mixed edges enter their set.
__global__ void k_edges_classify(EdgeArgs a)
{
    const long long NY1 = a.ny + 1, NZ1 = a.nz + 1;
    const long long ntot = (long long)(a.nx + 1) * NY1 * NZ1;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ntot) return;
    const int i = (int)(g / (NY1 * NZ1)), j = (int)((g / NZ1) % NY1), k = (int)(g % NZ1);
    uint32_t uniform;
    const unsigned long long key = edge_key(a, i, j, k, &uniform);
    if (uniform != 0xffffffffu) a.ID[(size_t)a.comp * ntot + g] = uniform;
    if (key == KEY_EMPTY) return;
    const unsigned long long index = (unsigned long long)a.comp * ntot + g;  // the reference's loop order
    unsigned int s = key_hash(key, a.cap);
    for (unsigned int probe = 0; probe < a.cap; ++probe, s = (s + 1) & (a.cap - 1)) {
        const unsigned long long prev = atomicCAS(&a.keys[s], KEY_EMPTY, key);
        if (prev == KEY_EMPTY || prev == key) {
            atomicMin(&a.first[s], index);
            return;
        }
    }
    atomicOr(a.overflow, 1u);
}

// Pass 2: mixed edges take the material the host resolved for their set.
__global__ void k_edges_scatter(EdgeArgs a)
{
    const long long NY1 = a.ny + 1, NZ1 = a.nz + 1;
    const long long ntot = (long long)(a.nx + 1) * NY1 * NZ1;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ntot) return;
    const int i = (int)(g / (NY1 * NZ1)), j = (int)((g / NZ1) % NY1), k = (int)(g % NZ1);
    uint32_t uniform;
    const unsigned long long key = edge_key(a, i, j, k, &uniform);
    if (key == KEY_EMPTY) return;
    unsigned int s = key_hash(key, a.cap);
    for (unsigned int probe = 0; probe < a.cap; ++probe, s = (s + 1) & (a.cap - 1)) {
        if (a.keys[s] == key) {
            a.ID[(size_t)a.comp * ntot + g] = a.resolved[s];
            return;
        }
        if (a.keys[s] == KEY_EMPTY) return;  // not reached: every mixed key was inserted
    }
}

std::string g_edge_err;

// ---------------------------------------------------------------- no-allocation variant
// gpx_edges_static keeps its hash tables in one file-scope device array per
// GPU instead of memory it allocates.  Layout (STATIC_CAP slots):
//   ukeys, ufirst  u64[STATIC_CAP]    sets not given by the caller (classify table)
//   kkeys          u64[STATIC_CAP]    sets the caller resolved (lookup table)
//   kids           u32[STATIC_CAP]    their materials
//   ckeys, cfirst  u64[STATIC_CAP]    the classify table compacted, table order
//   skeys          u64[STATIC_MAX]    staging of the caller's resolved sets
//   sids           u32[STATIC_MAX]
//   words          u32[64]            0: overflow, 1: duplicate key, 2: count
// Tables are at most half full (STATIC_MAX = STATIC_CAP / 2 sets).
constexpr unsigned int STATIC_CAP = 1u << 16;
constexpr unsigned int STATIC_MAX = STATIC_CAP / 2;
constexpr size_t SO_UKEYS = 0;
constexpr size_t SO_UFIRST = SO_UKEYS + (size_t)STATIC_CAP * 8;
constexpr size_t SO_KKEYS = SO_UFIRST + (size_t)STATIC_CAP * 8;
constexpr size_t SO_KIDS = SO_KKEYS + (size_t)STATIC_CAP * 8;
constexpr size_t SO_CKEYS = SO_KIDS + (size_t)STATIC_CAP * 4;
constexpr size_t SO_CFIRST = SO_CKEYS + (size_t)STATIC_CAP * 8;
constexpr size_t SO_SKEYS = SO_CFIRST + (size_t)STATIC_CAP * 8;
constexpr size_t SO_SIDS = SO_SKEYS + (size_t)STATIC_MAX * 8;
constexpr size_t SO_WORDS = SO_SIDS + (size_t)STATIC_MAX * 4;
constexpr size_t STATIC_BYTES = SO_WORDS + 256;

struct StaticTables {
    unsigned long long *ukeys, *ufirst, *kkeys, *ckeys, *cfirst, *skeys;
    uint32_t *kids, *sids, *words;
};

StaticTables static_tables(unsigned char *base)
{
    StaticTables t;
    t.ukeys = (unsigned long long *)(base + SO_UKEYS);
    t.ufirst = (unsigned long long *)(base + SO_UFIRST);
    t.kkeys = (unsigned long long *)(base + SO_KKEYS);
    t.kids = (uint32_t *)(base + SO_KIDS);
    t.ckeys = (unsigned long long *)(base + SO_CKEYS);
    t.cfirst = (unsigned long long *)(base + SO_CFIRST);
    t.skeys = (unsigned long long *)(base + SO_SKEYS);
    t.sids = (uint32_t *)(base + SO_SIDS);
    t.words = (uint32_t *)(base + SO_WORDS);
    return t;
}

// Empty tables, zero flags (a kernel, so the clearing is an ordinary launch).
__global__ void k_static_clear(StaticTables t)
{
    const unsigned int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < STATIC_CAP) {
        t.ukeys[g] = KEY_EMPTY;
        t.ufirst[g] = KEY_EMPTY;
        t.kkeys[g] = KEY_EMPTY;
        t.kids[g] = 0;
    }
    if (g < 64) t.words[g] = 0;
}

// The caller's resolved sets into the lookup table; a key given twice sets
// the duplicate flag.
__global__ void k_static_known(StaticTables t, int n)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const unsigned long long key = t.skeys[g];
    unsigned int s = key_hash(key, STATIC_CAP);
    for (unsigned int probe = 0; probe < STATIC_CAP; ++probe, s = (s + 1) & (STATIC_CAP - 1)) {
        const unsigned long long prev = atomicCAS(&t.kkeys[s], KEY_EMPTY, key);
        if (prev == KEY_EMPTY) {
            t.kids[s] = t.sids[g];
            return;
        }
        if (prev == key) {
            atomicOr(&t.words[1], 1u);
            return;
        }
    }
    atomicOr(&t.words[0], 1u);
}

// Uniform edges get their material, mixed edges of a resolved set get the
// set's material, the other mixed edges enter the classify table with their
// first edge.  The lookup table is complete before this kernel starts.
__global__ void k_edges_static(EdgeArgs a, StaticTables t)
{
    const long long NY1 = a.ny + 1, NZ1 = a.nz + 1;
    const long long ntot = (long long)(a.nx + 1) * NY1 * NZ1;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ntot) return;
    const int i = (int)(g / (NY1 * NZ1)), j = (int)((g / NZ1) % NY1), k = (int)(g % NZ1);
    uint32_t uniform;
    const unsigned long long key = edge_key(a, i, j, k, &uniform);
    if (uniform != 0xffffffffu) a.ID[(size_t)a.comp * ntot + g] = uniform;
    if (key == KEY_EMPTY) return;
    unsigned int s = key_hash(key, STATIC_CAP);
    for (unsigned int probe = 0; probe < STATIC_CAP; ++probe, s = (s + 1) & (STATIC_CAP - 1)) {
        const unsigned long long kk = t.kkeys[s];
        if (kk == key) {
            a.ID[(size_t)a.comp * ntot + g] = t.kids[s];
            return;
        }
        if (kk == KEY_EMPTY) break;
    }
    const unsigned long long index = (unsigned long long)a.comp * ntot + g;
    s = key_hash(key, STATIC_CAP);
    for (unsigned int probe = 0; probe < STATIC_MAX; ++probe, s = (s + 1) & (STATIC_CAP - 1)) {
        const unsigned long long prev = atomicCAS(&t.ukeys[s], KEY_EMPTY, key);
        if (prev == KEY_EMPTY || prev == key) {
            atomicMin(&t.ufirst[s], index);
            return;
        }
    }
    atomicOr(&t.words[0], 1u);
}

// The classify table's sets, compacted in table order (one block of 1024
// threads, 64 consecutive slots per thread); words[2] = their number.
__global__ void __launch_bounds__(1024) k_static_compact(StaticTables t)
{
    constexpr int PER = STATIC_CAP / 1024;
    __shared__ uint32_t part[1024];
    const int tid = threadIdx.x;
    const unsigned int s0 = (unsigned int)tid * PER;
    uint32_t n = 0;
    for (int q = 0; q < PER; ++q) n += t.ukeys[s0 + q] != KEY_EMPTY;
    part[tid] = n;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {  // inclusive scan
        const uint32_t v = tid >= d ? part[tid - d] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t o = part[tid] - n;
    for (int q = 0; q < PER; ++q) {
        const unsigned long long key = t.ukeys[s0 + q];
        if (key == KEY_EMPTY) continue;
        t.ckeys[o] = key;
        t.cfirst[o] = t.ufirst[s0 + q];
        ++o;
    }
    if (tid == 1023) t.words[2] = part[1023];
}

}  // namespace

struct gpx_edges {
    int device = 0;
    int nx = 0, ny = 0, nz = 0;
    long long ntot = 0, ncell = 0;
    hipStream_t stream = nullptr;
    uint32_t *solid = nullptr, *ID = nullptr;
    uint8_t *rigid[2] = {nullptr, nullptr};
    unsigned long long *keys = nullptr, *first = nullptr;
    uint32_t *resolved = nullptr;
    unsigned int *overflow = nullptr;
    unsigned int cap = 0;
    std::string err;
};

static int efail(gpx_edges *b, int code, const std::string &m)
{
    if (b) b->err = m;
    else g_edge_err = m;
    return code;
}

#define EDGE_TRY(call)                                                                           \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return efail(b, (int)e_, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" const char *gpx_edges_last_error(const gpx_edges *b)
{
    return b ? b->err.c_str() : g_edge_err.c_str();
}

extern "C" int gpx_edges_create(int device, int nx, int ny, int nz, const uint32_t *solid, const uint8_t *rigidE,
                                const uint8_t *rigidH, const uint32_t *ID, int capacity, gpx_edges **out)
{
    if (!out || !solid || !rigidE || !rigidH || !ID || nx < 1 || ny < 1 || nz < 1 || capacity < 1 || capacity > (1 << 28))
        return efail(nullptr, GPX_EARG, "gpx_edges_create: bad argument");
    *out = nullptr;
    gpx_edges *b = new gpx_edges();
    b->device = device;
    b->nx = nx; b->ny = ny; b->nz = nz;
    b->ncell = (long long)nx * ny * nz;
    b->ntot = (long long)(nx + 1) * (ny + 1) * (nz + 1);
    unsigned int cap = 1;
    while (cap < 2u * (unsigned int)capacity) cap <<= 1;
    b->cap = cap;
    auto clean = [&](int rc) {
        gpx_edges_destroy(b);
        return rc;
    };
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&b->solid, b->ncell * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&b->rigid[0], b->ncell);
    if (e == hipSuccess) e = hipMalloc((void **)&b->rigid[1], b->ncell);
    if (e == hipSuccess) e = hipMalloc((void **)&b->ID, 6 * b->ntot * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&b->keys, (size_t)cap * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&b->first, (size_t)cap * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&b->resolved, (size_t)cap * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&b->overflow, 4);
    if (e == hipSuccess) e = hipMemcpyAsync(b->solid, solid, b->ncell * 4, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->rigid[0], rigidE, b->ncell, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->rigid[1], rigidH, b->ncell, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->ID, ID, 6 * b->ntot * 4, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return clean(efail(nullptr, (int)e, std::string("gpx_edges_create: ") + hipGetErrorString(e)));
    *out = b;
    return 0;
}

extern "C" int gpx_edges_destroy(gpx_edges *b)
{
    if (!b) return 0;
    hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    for (void *p : {(void *)b->solid, (void *)b->rigid[0], (void *)b->rigid[1], (void *)b->ID, (void *)b->keys,
                    (void *)b->first, (void *)b->resolved, (void *)b->overflow})
        if (p) hipFree(p);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

static EdgeArgs edge_args(gpx_edges *b, int comp)
{
    EdgeArgs a;
    a.nx = b->nx; a.ny = b->ny; a.nz = b->nz;
    a.comp = comp;
    a.solid = b->solid;
    a.rigid = b->rigid[comp < 3 ? 0 : 1];
    a.ID = b->ID;
    a.keys = b->keys;
    a.first = b->first;
    a.cap = b->cap;
    a.overflow = b->overflow;
    a.resolved = b->resolved;
    return a;
}

// Classifies the edges of E (field 0) or H (field 1): uniform edges are
// written, and the distinct sets of the mixed ones are returned as (packed
// sorted materials, first edge index = comp * ntot + flat point index), at most
// `max` of them, in table order; *count is their number.
extern "C" int gpx_edges_classify(gpx_edges *b, int field, unsigned long long *keys, unsigned long long *first,
                                  int max, int *count)
{
    if (!b || (field != 0 && field != 1) || !count || max < 0) return efail(b, GPX_EARG, "gpx_edges_classify: bad argument");
    EDGE_TRY(hipSetDevice(b->device));
    EDGE_TRY(hipMemsetAsync(b->keys, 0xff, (size_t)b->cap * 8, b->stream));
    EDGE_TRY(hipMemsetAsync(b->first, 0xff, (size_t)b->cap * 8, b->stream));
    EDGE_TRY(hipMemsetAsync(b->overflow, 0, 4, b->stream));
    const int B = 256;
    const unsigned int grid = (unsigned int)((b->ntot + B - 1) / B);
    for (int c = 3 * field; c < 3 * field + 3; ++c) {
        hipLaunchKernelGGL(k_edges_classify, dim3(grid), dim3(B), 0, b->stream, edge_args(b, c));
        EDGE_TRY(hipGetLastError());
    }
    unsigned int overflow = 0;
    std::vector<unsigned long long> hk(b->cap), hf(b->cap);
    EDGE_TRY(hipMemcpyAsync(&overflow, b->overflow, 4, hipMemcpyDeviceToHost, b->stream));
    EDGE_TRY(hipMemcpyAsync(hk.data(), b->keys, (size_t)b->cap * 8, hipMemcpyDeviceToHost, b->stream));
    EDGE_TRY(hipMemcpyAsync(hf.data(), b->first, (size_t)b->cap * 8, hipMemcpyDeviceToHost, b->stream));
    EDGE_TRY(hipStreamSynchronize(b->stream));
    if (overflow)
        return efail(b, GPX_ELIMIT, "gpx_edges_classify: more distinct material sets than the table holds");
    int n = 0;
    for (unsigned int s = 0; s < b->cap; ++s) {
        if (hk[s] == KEY_EMPTY) continue;
        if (n < max) {
            if (keys) keys[n] = hk[s];
            if (first) first[n] = hf[s];
        }
        ++n;
    }
    *count = n;
    return 0;
}

// The material of each set returned by the last gpx_edges_classify of
// `field`, in the same order: written to every mixed edge of the set.
extern "C" int gpx_edges_resolve(gpx_edges *b, int field, const uint32_t *numids, int count)
{
    if (!b || (field != 0 && field != 1) || count < 0 || (count && !numids))
        return efail(b, GPX_EARG, "gpx_edges_resolve: bad argument");
    EDGE_TRY(hipSetDevice(b->device));
    std::vector<unsigned long long> hk(b->cap);
    EDGE_TRY(hipMemcpyAsync(hk.data(), b->keys, (size_t)b->cap * 8, hipMemcpyDeviceToHost, b->stream));
    EDGE_TRY(hipStreamSynchronize(b->stream));
    std::vector<uint32_t> table(b->cap, 0);
    int n = 0;
    for (unsigned int s = 0; s < b->cap; ++s) {
        if (hk[s] == KEY_EMPTY) continue;
        if (n >= count) return efail(b, GPX_EARG, "gpx_edges_resolve: fewer materials than sets");
        table[s] = numids[n++];
    }
    if (n != count) return efail(b, GPX_EARG, "gpx_edges_resolve: more materials than sets");
    EDGE_TRY(hipMemcpyAsync(b->resolved, table.data(), (size_t)b->cap * 4, hipMemcpyHostToDevice, b->stream));
    const int B = 256;
    const unsigned int grid = (unsigned int)((b->ntot + B - 1) / B);
    for (int c = 3 * field; c < 3 * field + 3; ++c) {
        hipLaunchKernelGGL(k_edges_scatter, dim3(grid), dim3(B), 0, b->stream, edge_args(b, c));
        EDGE_TRY(hipGetLastError());
    }
    EDGE_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

extern "C" int gpx_edges_download(gpx_edges *b, uint32_t *ID)
{
    if (!b || !ID) return efail(b, GPX_EARG, "gpx_edges_download: bad argument");
    EDGE_TRY(hipSetDevice(b->device));
    EDGE_TRY(hipMemcpyAsync(ID, b->ID, 6 * b->ntot * 4, hipMemcpyDeviceToHost, b->stream));
    EDGE_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

// ---------------------------------------------------------------- no-allocation entry point
// The device array behind gpx_edges_static (3.1 MiB; each GPU that runs it
// holds its own copy until the process ends).  Host code never passes this
// name to a kernel or a copy (on the host it is only a stand-in for the
// symbol); it uses the device address hipGetSymbolAddress gives for the
// current GPU.
static __device__ __align__(256) unsigned char scratch[STATIC_BYTES];

namespace {
constexpr int STATIC_DEVICES = 64;
std::mutex g_static_mu;
unsigned char *g_static_ptr[STATIC_DEVICES] = {nullptr};

// The device address of `scratch` on `device`, looked up once per device with
// that device current.
int static_base(int device, unsigned char **out)
{
    std::lock_guard<std::mutex> lock(g_static_mu);
    if (!g_static_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return efail(nullptr, e != hipSuccess ? (int)e : GPX_EARG,
                         std::string("gpx_edges_static: hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_static_ptr[device] = (unsigned char *)p;
    }
    *out = g_static_ptr[device];
    return 0;
}
}  // namespace

extern "C" int gpx_edges_static(int device, int nx, int ny, int nz, const uint32_t *solid, const uint8_t *rigid,
                                uint32_t *ID, int field, int nknown, const unsigned long long *known_keys,
                                const uint32_t *known_ids, unsigned long long *keys, unsigned long long *first,
                                int max, int *count)
{
    gpx_edges *b = nullptr;  // errors go to gpx_edges_last_error(NULL)
    if (device < 0 || device >= STATIC_DEVICES || nx < 1 || ny < 1 || nz < 1 || !solid || !rigid || !ID ||
        (field != 0 && field != 1) || nknown < 0 || (nknown && (!known_keys || !known_ids)) || max < 0 || !count)
        return efail(b, GPX_EARG, "gpx_edges_static: bad argument");
    if (nknown > (int)STATIC_MAX)
        return efail(b, GPX_ELIMIT, "gpx_edges_static: at most " + std::to_string(STATIC_MAX) +
                                         " resolved material sets per call (" + std::to_string(nknown) + " given)");
    EDGE_TRY(hipSetDevice(device));
    unsigned char *base = nullptr;
    if (int rc = static_base(device, &base)) return rc;
    const StaticTables t = static_tables(base);
    hipStream_t s = nullptr;  // the default stream: ordered after the caller's work on it
    hipLaunchKernelGGL(k_static_clear, dim3(STATIC_CAP / 256), dim3(256), 0, s, t);
    EDGE_TRY(hipGetLastError());
    if (nknown) {
        EDGE_TRY(hipMemcpyAsync(t.skeys, known_keys, (size_t)nknown * 8, hipMemcpyHostToDevice, s));
        EDGE_TRY(hipMemcpyAsync(t.sids, known_ids, (size_t)nknown * 4, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_static_known, dim3((nknown + 255) / 256), dim3(256), 0, s, t, nknown);
        EDGE_TRY(hipGetLastError());
    }
    EdgeArgs a;
    a.nx = nx; a.ny = ny; a.nz = nz;
    a.solid = solid;
    a.rigid = rigid;
    a.ID = ID;
    a.keys = nullptr; a.first = nullptr; a.cap = 0; a.overflow = nullptr; a.resolved = nullptr;
    const long long ntot = (long long)(nx + 1) * (ny + 1) * (nz + 1);
    const int B = 256;
    const unsigned int grid = (unsigned int)((ntot + B - 1) / B);
    for (int c = 3 * field; c < 3 * field + 3; ++c) {
        a.comp = c;
        hipLaunchKernelGGL(k_edges_static, dim3(grid), dim3(B), 0, s, a, t);
        EDGE_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_static_compact, dim3(1), dim3(1024), 0, s, t);
    EDGE_TRY(hipGetLastError());
    uint32_t words[3] = {0, 0, 0};
    EDGE_TRY(hipMemcpyAsync(words, t.words, sizeof(words), hipMemcpyDeviceToHost, s));
    EDGE_TRY(hipStreamSynchronize(s));
    if (words[1]) return efail(b, GPX_EARG, "gpx_edges_static: a resolved material set is given twice");
    if (words[0])
        return efail(b, GPX_ELIMIT, "gpx_edges_static: more than " + std::to_string(STATIC_MAX) +
                                         " distinct material sets to resolve (the table is full)");
    const int n = (int)words[2];
    if (n > (int)STATIC_MAX)
        return efail(b, GPX_ELIMIT, "gpx_edges_static: more than " + std::to_string(STATIC_MAX) +
                                         " distinct material sets to resolve (" + std::to_string(n) + ")");
    const int m = n < max ? n : max;
    if (m > 0) {
        if (keys) EDGE_TRY(hipMemcpyAsync(keys, t.ckeys, (size_t)m * 8, hipMemcpyDeviceToHost, s));
        if (first) EDGE_TRY(hipMemcpyAsync(first, t.cfirst, (size_t)m * 8, hipMemcpyDeviceToHost, s));
        EDGE_TRY(hipStreamSynchronize(s));
    }
    *count = n;
    return 0;
}
