// C ABI of the GPU FOD segmentation (include/mrfixel.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mrfixel.h"
#include "capi_common.h"
#include "host/fmls.h"
#include "host/sh.h"
#include "kernels/fmls.h"

namespace {

// Voxels per pass: one lane each, 2048 wavefronts (two per SIMD, the segmentation kernel's
// occupancy); the work arrays take about 110 KB per voxel at 1281 directions (14 GB for a full
// pass), sized for the worst case so that no voxel is deferred, and a pass is cut down to what
// half the free device memory holds.  (The CPU emulation build of the kernels sets a small chunk
// to run several passes.)
#ifndef MRFIXEL_CHUNK
#define MRFIXEL_CHUNK 131072
#endif
constexpr size_t kChunk = MRFIXEL_CHUNK;

const mrcsd::fmls::DirectionSet& direction_set(const double* az_el, int n, mrcsd::fmls::DirectionSet& own) {
  if (!az_el) return mrcsd::fmls::tessellation_1281();
  own = mrcsd::fmls::make_direction_set(std::vector<double>(az_el, az_el + 2 * size_t(n)));
  return own;
}

}  // namespace

struct mrfixel_ctx {
  int device = 0;
  int n_sh = 0, lmax = 0, ndir = 0;
  mrcsd::FmlsDeviceSet S{};
  mrcsd::DeviceBuffer<double> unit, weights, sht, al;
  mrcsd::DeviceBuffer<int16_t> adj, cell_dirs;
  mrcsd::DeviceBuffer<int> cell_start;
  mrcsd::DeviceBuffer<unsigned char> work;   // fmls_work_bytes(ndir, chunk) bytes
  size_t chunk = 0;                           // slots of the work arrays
  mrcsd::DeviceBuffer<unsigned long long> sums;
};

namespace mrcsd {

namespace {
size_t align256(size_t n) { return (n + 255) & ~size_t(255); }
}  // namespace

size_t fmls_work_bytes(int ndir, size_t chunk) {
  const size_t n = size_t(ndir) * chunk;
  // doubles: amps (fmls_amp_ld(ndir) per slot), peak, integ (n each), mean, rev (3 n each); int16:
  // lab, nxt, head, tail, np; uint32: ordm (n); int32: npos, nlobes (one per slot); uint32: items
  // (n); the item count -- the pieces fmls_work_carve takes, in its order
  return align256(8 * size_t(fmls_amp_ld(ndir)) * chunk) + 2 * align256(8 * n) + 2 * align256(8 * 3 * n) +
         5 * align256(2 * n) + align256(4 * n) + 2 * align256(4 * chunk) + align256(4 * n) + align256(8);
}

FmlsWork fmls_work_carve(void* base, int ndir, size_t chunk) {
  const size_t n = size_t(ndir) * chunk;
  unsigned char* p = static_cast<unsigned char*>(base);
  auto take = [&](size_t bytes) {
    unsigned char* q = p;
    p += align256(bytes);
    return q;
  };
  FmlsWork W{};
  W.amps = reinterpret_cast<double*>(take(8 * size_t(fmls_amp_ld(ndir)) * chunk));
  W.peak = reinterpret_cast<double*>(take(8 * n));
  W.integ = reinterpret_cast<double*>(take(8 * n));
  W.mean = reinterpret_cast<double*>(take(8 * 3 * n));
  W.rev = reinterpret_cast<double*>(take(8 * 3 * n));
  W.lab = reinterpret_cast<int16_t*>(take(2 * n));
  W.nxt = reinterpret_cast<int16_t*>(take(2 * n));
  W.head = reinterpret_cast<int16_t*>(take(2 * n));
  W.tail = reinterpret_cast<int16_t*>(take(2 * n));
  W.np = reinterpret_cast<int16_t*>(take(2 * n));
  W.ordm = reinterpret_cast<uint32_t*>(take(4 * n));
  W.npos = reinterpret_cast<int32_t*>(take(4 * chunk));
  W.nlobes = reinterpret_cast<int32_t*>(take(4 * chunk));
  W.items = reinterpret_cast<uint32_t*>(take(4 * n));
  W.nitems = reinterpret_cast<unsigned long long*>(take(8));
  return W;
}

}  // namespace mrcsd

extern "C" {

void mrfixel_options_default(mrfixel_options* opt) {
  if (!opt) return;
  std::memset(opt, 0, sizeof(*opt));
  opt->integral_threshold = 0.0;     // FMLS_INTEGRAL_THRESHOLD_DEFAULT (fmls.h:31)
  opt->peak_value_threshold = 0.1;   // FMLS_PEAK_VALUE_THRESHOLD_DEFAULT (fmls.h:32)
  opt->lobe_merge_ratio = 1.0;       // FMLS_MERGE_RATIO_BRIDGE_TO_PEAK_DEFAULT (fmls.h:33)
}

int mrfixel_direction_set(const double* az_el, int n, mrfixel_set_info* info, double* unit, int* adj, double* weights,
                          int* cell_start, int* cell_dirs) {
  return mrcsd::guarded([&]() -> int {
    if (!info) return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_direction_set: null info");
    if (az_el && (n < 64 || n > mrcsd::kFmlsMaxDirs))
      return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_direction_set: 64 to 1536 directions");
    mrcsd::fmls::DirectionSet own;
    const auto& s = direction_set(az_el, n, own);
    info->ndir = s.n;
    info->n_az = s.n_az;
    info->n_el = s.n_el;
    info->n_cell_dirs = int(s.cell_dirs.size());
    if (unit) std::memcpy(unit, s.unit.data(), s.unit.size() * sizeof(double));
    if (adj) std::memcpy(adj, s.adj.data(), s.adj.size() * sizeof(int));
    if (weights) std::memcpy(weights, s.weights.data(), s.weights.size() * sizeof(double));
    if (cell_start) std::memcpy(cell_start, s.cell_start.data(), s.cell_start.size() * sizeof(int));
    if (cell_dirs) std::memcpy(cell_dirs, s.cell_dirs.data(), s.cell_dirs.size() * sizeof(int));
    return MRCSD_OK;
  });
}

int mrfixel_legendre_table(int lmax, int nrows, double* out) {
  return mrcsd::guarded([&]() -> int {
    if (!out || lmax < 0 || lmax > 16 || (lmax & 1) || nrows < 2)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_legendre_table: even lmax up to 16, at least 2 rows");
    const auto t = mrcsd::fmls::legendre_table(lmax, nrows);
    std::memcpy(out, t.data(), t.size() * sizeof(double));
    return MRCSD_OK;
  });
}

int mrfixel_create(const mrfixel_options* opt, mrfixel_ctx** out) {
  return mrcsd::guarded([&]() -> int {
    if (!opt || !out) return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_create: null argument");
    *out = nullptr;
    const int lmax = mrcsd::sh::LforN(opt->n_sh);
    if (opt->n_sh < 1 || lmax < 2 || lmax > mrcsd::kFmlsMaxLmax || mrcsd::sh::NforL(lmax) != opt->n_sh)
      return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_create: the FOD must be an even SH series of lmax 2 to 12 (n_sh " +
                                            std::to_string(opt->n_sh) + ")");
    if (opt->directions && (opt->n_directions < 64 || opt->n_directions > mrcsd::kFmlsMaxDirs))
      return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_create: 64 to 1536 segmentation directions");
    if (opt->maxnum < 0) return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_create: maxnum must be positive (or 0: no limit)");
    if (!(opt->lobe_merge_ratio >= 0.0)) return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_create: invalid lobe merge ratio");
    int rc = mrcsd::require_gpu();
    if (rc) return rc;
    mrcsd::fmls::DirectionSet own;
    const auto& s = direction_set(opt->directions, opt->n_directions, own);
    auto ctx = std::make_unique<mrfixel_ctx>();
    MRCSD_HIP(hipGetDevice(&ctx->device));
    ctx->n_sh = opt->n_sh;
    ctx->lmax = lmax;
    ctx->ndir = s.n;
    // the amplitude transform of the directions as the Segmenter builds it (fmls.cpp:150-157):
    // azimuth and elevation taken back from the unit vectors
    mrcsd::Matrix azel(s.n, 2);
    for (int i = 0; i < s.n; ++i) {
      azel(i, 0) = std::atan2(s.unit[size_t(3 * i + 1)], s.unit[size_t(3 * i)]);
      azel(i, 1) = std::acos(s.unit[size_t(3 * i + 2)]);
    }
    const mrcsd::Matrix T = mrcsd::sh::transform(azel, lmax);
    std::vector<int16_t> adj16(s.adj.begin(), s.adj.end()), cells16(s.cell_dirs.begin(), s.cell_dirs.end());
    const int al_rows = 2 * s.n;   // PrecomputedAL (lmax, 2 * dirs.size()) (fmls.cpp:146)
    const auto al = mrcsd::fmls::legendre_table(lmax, al_rows);
    ctx->unit.upload(s.unit.data(), s.unit.size());
    ctx->weights.upload(s.weights.data(), s.weights.size());
    ctx->sht.upload(T.a.data(), T.a.size());
    ctx->al.upload(al.data(), al.size());
    ctx->adj.upload(adj16.data(), adj16.size());
    ctx->cell_dirs.upload(cells16.data(), cells16.size());
    ctx->cell_start.upload(s.cell_start.data(), s.cell_start.size());
    auto& S = ctx->S;
    S.ndir = s.n;
    S.unit = ctx->unit.get();
    S.adj = ctx->adj.get();
    S.weights = ctx->weights.get();
    S.sht = ctx->sht.get();
    S.nsh = opt->n_sh;
    S.lmax = lmax;
    S.n_az = s.n_az;
    S.n_el = s.n_el;
    S.az_step = s.az_step;
    S.el_step = s.el_step;
    S.az_begin = s.az_begin;
    S.el_begin = s.el_begin;
    S.cell_start = ctx->cell_start.get();
    S.cell_dirs = ctx->cell_dirs.get();
    S.al = ctx->al.get();
    S.al_ndir = al_rows;
    S.integral_threshold = opt->integral_threshold;
    S.peak_threshold = opt->peak_value_threshold;
    S.merge_ratio = opt->lobe_merge_ratio;
    S.maxnum = opt->maxnum;
    S.dir_from_peak = opt->dir_from_peak ? 1 : 0;
    int lds_cu = 0, lds_block = 0;
    MRCSD_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, ctx->device));
    MRCSD_HIP(hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
    S.lanes8 = mrcsd::fmls_grow_lanes(s.n, 1, lds_cu, lds_block);
    S.lanes16 = mrcsd::fmls_grow_lanes(s.n, 2, lds_cu, lds_block);
    if (S.lanes16 < 1)
      return mrcsd::fail(MRCSD_ERR_NOGPU, "mrfixel_create: the device's LDS cannot hold a voxel's lobe labels");
    *out = ctx.release();
    return MRCSD_OK;
  });
}

void mrfixel_destroy(mrfixel_ctx* ctx) { delete ctx; }

int mrfixel_segment(mrfixel_ctx* ctx, const float* fod, int ld_fod, const uint8_t* mask, size_t nvox, uint32_t* count,
                    uint32_t* offset, float* fixels, size_t capacity, float* summary, uint64_t* total, void* stream) {
  return mrcsd::guarded([&]() -> int {
    if (!ctx) return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_segment: null context");
    if (!total) return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_segment: null total");
    if (nvox && (!fod || !count || !offset || (capacity && !fixels)))
      return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_segment: null fod, count, offset or fixels");
    if (ld_fod < ctx->n_sh) return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_segment: ld_fod smaller than n_sh");
    if (nvox > 0xFFFFFFFFull) return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_segment: at most 2^32 - 1 voxels per call");
    MRCSD_HIP(hipSetDevice(ctx->device));
    const hipStream_t st = static_cast<hipStream_t>(stream);
    // slots of the work arrays: this call's voxels (whole blocks) up to a full chunk; a larger set
    // kept from an earlier call is used as it is
    size_t want = std::min(kChunk, std::max<size_t>(size_t(mrcsd::kFmlsBlock),
                                                    (nvox + mrcsd::kFmlsBlock - 1) / mrcsd::kFmlsBlock *
                                                        mrcsd::kFmlsBlock));
    if (ctx->chunk < want) {
      MRCSD_HIP(hipStreamSynchronize(st));   // a previous call may still use them
      ctx->work.reset();
      ctx->sums.reset();
      ctx->chunk = 0;
      size_t free_b = 0, total_b = 0;
      MRCSD_HIP(hipMemGetInfo(&free_b, &total_b));
      while (want > size_t(mrcsd::kFmlsBlock) && mrcsd::fmls_work_bytes(ctx->ndir, want) > free_b / 2)
        want = std::max(size_t(mrcsd::kFmlsBlock), want / 2 / mrcsd::kFmlsBlock * mrcsd::kFmlsBlock);
      ctx->work.alloc(mrcsd::fmls_work_bytes(ctx->ndir, want));
      ctx->sums.alloc(want / mrcsd::kFmlsBlock + 1);
      ctx->chunk = want;
    }
    mrcsd::FmlsLaunch L{};
    L.fod = fod;
    L.ld_fod = ld_fod;
    L.valid = mask;
    L.nvox = nvox;
    L.chunk = ctx->chunk;
    L.work = mrcsd::fmls_work_carve(ctx->work.get(), ctx->ndir, ctx->chunk);
    L.count = count;
    L.offset = offset;
    L.fixels = fixels;
    L.capacity = capacity;
    L.total = reinterpret_cast<unsigned long long*>(total);
    L.summary = summary;
    L.sums = ctx->sums.get();
    MRCSD_HIP(mrcsd::fmls_launch(ctx->S, L, st));
    return MRCSD_OK;
  });
}

int mrfixel_segment_host(mrfixel_ctx* ctx, const float* fod, int ld_fod, const uint8_t* mask, size_t nvox,
                         uint32_t* count, uint32_t* offset, float* fixels, size_t capacity, float* summary,
                         uint64_t* total) {
  return mrcsd::guarded([&]() -> int {
    if (!ctx) return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_segment_host: null context");
    if (!total) return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_segment_host: null total");
    if (nvox && (!fod || !count || !offset || (capacity && !fixels)))
      return mrcsd::fail(MRCSD_ERR_ARG, "mrfixel_segment_host: null fod, count, offset or fixels");
    MRCSD_HIP(hipSetDevice(ctx->device));
    mrcsd::DeviceBuffer<float> d_fod, d_fix, d_sum;
    mrcsd::DeviceBuffer<uint8_t> d_mask;
    mrcsd::DeviceBuffer<uint32_t> d_count, d_off;
    mrcsd::DeviceBuffer<unsigned long long> d_total;
    d_fod.upload(fod, nvox * size_t(ld_fod));
    if (mask) d_mask.upload(mask, nvox);
    d_count.alloc(nvox);
    d_off.alloc(nvox);
    d_fix.alloc(5 * capacity);
    if (summary) d_sum.alloc(5 * nvox);
    d_total.alloc(1);
    int rc = mrfixel_segment(ctx, d_fod.get(), ld_fod, mask ? d_mask.get() : nullptr, nvox, d_count.get(), d_off.get(),
                             d_fix.get(), capacity, summary ? d_sum.get() : nullptr,
                             reinterpret_cast<uint64_t*>(d_total.get()), nullptr);
    if (rc) return rc;
    MRCSD_HIP(hipDeviceSynchronize());
    unsigned long long t = 0;
    d_total.download(&t, 1);
    *total = t;
    d_count.download(count, nvox);
    d_off.download(offset, nvox);
    const size_t nfix = t < capacity ? size_t(t) : capacity;
    d_fix.download(fixels, 5 * nfix);
    if (summary) d_sum.download(summary, 5 * nvox);
    return MRCSD_OK;
  });
}

}  // extern "C"
