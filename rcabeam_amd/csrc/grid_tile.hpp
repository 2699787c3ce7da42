// Tile-level machinery of the Cartesian-grid delay-and-sum kernels
// (das_mfma.hip: MFMA interpolation of 16-bit channels; das_grid.hip: VALU
// interpolation of complex64 channels).
//
// A workgroup owns an 8 x 8 lateral patch of voxels at one depth.  On a grid
// the delay separates: the receive term of RC depends on (x, z, channel) and
// its transmit term on (y, z, angle); CR swaps x and y.  A tile therefore
// needs 8 receive terms per channel and 8 transmit terms per angle (float64,
// split into integer sample, float32 fraction and unit phasor, common.hpp).
// Transmit terms are formed once per tile into LDS; every wave forms the
// receive terms of 8 consecutive channels at once (lane l: channel l / 8,
// position l % 8) and hands them out channel by channel.
// The first tap row is monotonic in both terms, so the row window of any set
// of voxels is exactly [first row of (min receive, min transmit), last row of
// (max receive, max transmit)] clipped to the closed sample window: the
// windows of the tile (staging) and of its 4 x 4 quarters come from min / max
// over 8 or 4 terms, with no tap table.  Channels whose dynamic aperture
// misses the whole tile are skipped (one bit per channel).
#pragma once

#include "ensemble.hpp"
#include "tile_staging.hpp"

namespace rcb {

constexpr int kGridPatch = 8;      // lateral voxels per tile side
constexpr int kGridMaxAngles = 16;
// Tile order: blocks of 4 depths x 4 y-tiles x 4 x-tiles (64 tiles, the
// workgroups one XCD runs at once), blocks depth-fastest.  Neighbouring tiles
// share receive terms along one axis and transmit terms along the other, so a
// block fetches fewer distinct sample rows into its XCD's L2 than 64
// consecutive depths of one lateral tile do (555 GB against 724 GB of L2
// misses per full-size float32 launch).
constexpr int kOrderZ = 4, kOrderY = 4, kOrderX = 4;
constexpr int kOrderBlock = kOrderZ * kOrderY * kOrderX;

// A delay term as an ordered key: integer sample, then fraction (a float in
// [0, 1), whose bits order like the value).  Inactive terms carry the
// extreme key of the reduction they take part in.
struct Key {
  int i;
  uint32_t f;
};
__device__ __forceinline__ bool key_less(Key a, Key b) { return a.i < b.i || (a.i == b.i && a.f < b.f); }
__device__ __forceinline__ Key key_min(Key a, Key b) { return key_less(b, a) ? b : a; }
__device__ __forceinline__ Key key_max(Key a, Key b) { return key_less(a, b) ? b : a; }
constexpr Key kKeyHigh{0x7fffffff, 0u};
constexpr Key kKeyLow{-0x7fffffff - 1, 0u};

template <int CTRL>
__device__ __forceinline__ Key dpp_key(Key k) {
  return Key{__builtin_amdgcn_update_dpp(0, k.i, CTRL, 0xf, 0xf, false),
             static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(k.f), CTRL, 0xf, 0xf, false))};
}

// First and last sample rows of the tap of receive key r and transmit key t,
// computed as combine() computes the tap (integer sum, float32 fraction sum,
// carry); x = first row, y = last row (first + 1 unless the fraction is 0).
__device__ __forceinline__ int2 tap_rows(Key r, Key t) {
  int i = r.i + t.i;
  float f = __uint_as_float(r.f) + __uint_as_float(t.f);
  if (f >= 1.0f) {
    i += 1;
    f -= 1.0f;
  }
  return make_int2(i, f != 0.0f ? i + 1 : i);
}

// The tap of receive term r and transmit term t exactly as combine() forms
// it, without branches: first row (-1: no contribution), weight, phasor.
__device__ __forceinline__ Tap tap_of(const Term& r, const Term& t, int ns) {
  int i = r.i + t.i;
  float f = r.f + t.f;
  const bool carry = f >= 1.0f;
  i += carry ? 1 : 0;
  f = carry ? f - 1.0f : f;
  const bool valid =
      r.i != kInactive && t.i != kInactive && i >= 0 && (i < ns - 1 || (i == ns - 1 && f == 0.0f));
  Tap out;
  out.i0 = valid ? i : -1;
  out.w = valid ? f : 0.f;
  out.c = valid ? r.c * t.c - r.s * t.s : 0.f;
  out.s = valid ? r.c * t.s + r.s * t.c : 0.f;
  return out;
}

// Window of the voxels whose receive terms span [rlo, rhi] and transmit terms
// [tlo, thi]: rows of valid taps (closed window 0 <= s <= ns - 1).  Empty
// (x > y) if either set is empty.
__device__ __forceinline__ int2 window(Key rlo, Key rhi, Key tlo, Key thi, int ns) {
  if (rlo.i == kKeyHigh.i || tlo.i == kKeyHigh.i) return make_int2(kBig, -kBig);
  const int lo = tap_rows(rlo, tlo).x;
  const int hi = tap_rows(rhi, thi).y;
  return make_int2(max(lo, 0), min(hi, ns - 1));
}

struct GridArgs {
  const void* rc;  // packed (channels, angles, samples, frames) RC channels
  const void* cr;
  const float* x;
  const float* z;
  const float* y;
  const float* xe;
  const float* ye;
  const float* angles;
  const float* starts;
  float* pd;
  void* signal;
  long long tiles;  // tile slots: whole order blocks, some of them outside the grid
  int nx, nz, ny, nbx, nby, nzb, nyb;  // tiles along x, y; order blocks along z, y
  int ns, nc, na, nt;
  int max_rows, out_storage, vec, stride;
  double c, fs, fd, fn;
  const unsigned* absmax;  // split FP16 channels: bits of the joint max |component| (scale exponent)
  int q_begin, q_end;      // channels [q_begin, q_end) of the 2 nc (RC then CR) take part
  int accumulate;          // VALU kernel, complex64 signal: add the sums to the signal already there
};

// Grid arguments from the ensemble arguments (pd, signal and channels as
// given; tiles of 8 x 8 lateral voxels per depth).
inline GridArgs make_grid_args(const EnsembleArgs& e, const float* x, const float* z, const float* y, int nx, int nz,
                               int ny, int out_storage) {
  GridArgs a{};
  a.rc = e.rc;
  a.cr = e.cr;
  a.x = x;
  a.z = z;
  a.y = y;
  a.xe = e.xe;
  a.ye = e.ye;
  a.angles = e.angles;
  a.starts = e.starts;
  a.pd = e.pd;
  a.signal = e.signal;
  a.nx = nx;
  a.nz = nz;
  a.ny = ny;
  a.nbx = (nx + kGridPatch - 1) / kGridPatch;
  a.nby = (ny + kGridPatch - 1) / kGridPatch;
  a.nzb = (nz + kOrderZ - 1) / kOrderZ;
  a.nyb = (a.nby + kOrderY - 1) / kOrderY;
  const long long nxb = (a.nbx + kOrderX - 1) / kOrderX;
  a.tiles = static_cast<long long>(a.nzb) * a.nyb * nxb * kOrderBlock;
  a.ns = e.ns;
  a.nc = e.nc;
  a.na = e.na;
  a.nt = e.nt;
  a.out_storage = out_storage;
  a.q_begin = 0;
  a.q_end = 2 * e.nc;
  a.accumulate = 0;
  a.c = e.c;
  a.fs = e.fs;
  a.fd = e.fd;
  a.fn = e.fn;
  return a;
}

// Tile-constant LDS tables: txt Term[2][16][8] transmit terms (aperture,
// angle, tile position), txw Key[2][16][6] their (min, max) over the 8
// positions, positions 0-3 and positions 4-7, live: one bit per channel.
constexpr int kTxtBytes = 2 * kGridMaxAngles * 8 * 16;
constexpr int kTxwBytes = 2 * kGridMaxAngles * 6 * 8;
__host__ __device__ inline int live_bytes(int nc) { return ((2 * nc + 63) / 64) * 8; }
__host__ __device__ inline int tile_table_bytes(int nc) { return kTxtBytes + kTxwBytes + live_bytes(nc); }

// What a wave knows about one channel: the receive term of each lane's own
// voxel, the tile window of every angle (lane a) and the window of every
// angle for one 4 x 4 quarter of the tile (lane a).
struct Channel {
  int q, cfg;
  Term rx;
  Windows w;  // tile windows
  Windows g;  // quarter windows
};

// Tile slot t (order blocks, above) placed XCD-aware: block b runs on XCD
// b % 8, so each XCD takes a contiguous run of slots.  Slots outside the grid
// are invalid.
template <int THREADS>
struct GridTile {
  const GridArgs& p;
  bool ok;
  int iz, x0, y0;
  double zv;
  Term* txt;
  Key* txw;
  uint64_t* live;

  __device__ GridTile(const GridArgs& args, unsigned char* tables) : p(args) {
    const long long per_xcd = (p.tiles + 7) / 8;
    const long long t = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    const int local = static_cast<int>(t % kOrderBlock);
    const long long block = t / kOrderBlock;
    const long long bz = block % p.nzb, rest = block / p.nzb;
    iz = static_cast<int>(bz * kOrderZ + local % kOrderZ);
    const int by = static_cast<int>((rest % p.nyb) * kOrderY + (local / kOrderZ) % kOrderY);
    const int bx = static_cast<int>((rest / p.nyb) * kOrderX + local / (kOrderZ * kOrderY));
    ok = t < p.tiles && iz < p.nz && by < p.nby && bx < p.nbx;
    if (!ok) iz = 0;
    x0 = bx * kGridPatch;
    y0 = by * kGridPatch;
    zv = ok ? static_cast<double>(p.z[iz]) : 0.0;
    txt = reinterpret_cast<Term*>(tables);
    txw = reinterpret_cast<Key*>(tables + kTxtBytes);
    live = reinterpret_cast<uint64_t*>(tables + kTxtBytes + kTxwBytes);
  }
  __device__ bool valid() const { return ok; }
  // axis_x: position along x (else y) of the 8 x 8 patch
  __device__ bool in_grid(bool axis_x, int pos) const { return axis_x ? x0 + pos < p.nx : y0 + pos < p.ny; }
  __device__ double coord(bool axis_x, int pos) const {
    return static_cast<double>(axis_x ? p.x[x0 + pos] : p.y[y0 + pos]);
  }
  __device__ long long voxel(int ix, int iy) const {  // tile-local indices; -1 outside the grid
    if (x0 + ix >= p.nx || y0 + iy >= p.ny) return -1;
    return (static_cast<long long>(x0 + ix) * p.nz + iz) * p.ny + (y0 + iy);
  }

  // Transmit terms, their min / max and the live-channel bits; ends with a
  // barrier, after which the tables are visible to every wave.
  __device__ void build_tables() {
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = uniform(tid / kWave);
    if (tid < 2 * kGridMaxAngles * 8) {
      const int cfg = tid >> 7, a = (tid >> 3) & 15, pos = tid & 7;
      const bool axis_x = cfg != 0;  // RC transmits along y, CR along x
      Term t{kInactive, 0.f, 0.f, 0.f};
      if (a < p.na && in_grid(axis_x, pos)) {
        double sn, cs;
        sincos(static_cast<double>(p.angles[a]), &sn, &cs);
        t = transmit_term(coord(axis_x, pos), zv, cs, sn, static_cast<double>(p.starts[a]), p.c, p.fs, p.fd);
      }
      txt[tid] = t;
    }
    const int nq = 2 * p.nc;
    for (int q0 = 0; q0 < nq; q0 += THREADS) {
      const int q = q0 + tid;
      bool on = false;
      if (q < nq) {
        const int cfg = q >= p.nc ? 1 : 0;
        const double e = static_cast<double>((cfg == 0 ? p.xe : p.ye)[q - cfg * p.nc]);
        for (int pos = 0; pos < kGridPatch; ++pos) {
          if (!in_grid(cfg == 0, pos)) continue;
          const double dv = coord(cfg == 0, pos) - e;  // same test as receive_term
          on = on || !(p.fn > 0.0) || fabs(dv) <= zv / (2.0 * p.fn);
        }
        on = on && q >= p.q_begin && q < p.q_end;
      }
      const uint64_t mask = __ballot(on);
      if (lane == 0 && q0 + kWave * wave < nq) live[(q0 >> 6) + wave] = mask;
    }
    __syncthreads();
    if (tid < 2 * kGridMaxAngles) {
      const Term* t = txt + tid * 8;
      Key lo[3] = {kKeyHigh, kKeyHigh, kKeyHigh}, hi[3] = {kKeyLow, kKeyLow, kKeyLow};
      for (int pos = 0; pos < kGridPatch; ++pos) {
        if (t[pos].i == kInactive) continue;
        const Key k{t[pos].i, __float_as_uint(t[pos].f)};
        lo[0] = key_min(lo[0], k);
        hi[0] = key_max(hi[0], k);
        lo[1 + (pos >> 2)] = key_min(lo[1 + (pos >> 2)], k);
        hi[1 + (pos >> 2)] = key_max(hi[1 + (pos >> 2)], k);
      }
      for (int k = 0; k < 3; ++k) {
        txw[tid * 6 + 2 * k] = lo[k];
        txw[tid * 6 + 2 * k + 1] = hi[k];
      }
    }
    __syncthreads();
  }

  // First live channel after q (2 nc if none); uniform.
  __device__ int next_live(int q) const {
    const int nq = 2 * p.nc;
    for (int word = (q + 1) >> 6; word * 64 < nq; ++word) {
      uint64_t m = live[word];
      if (word == (q + 1) >> 6) m &= ~0ull << ((q + 1) & 63);
      m = (static_cast<uint64_t>(uniform(static_cast<int>(m >> 32))) << 32) |
          static_cast<uint32_t>(uniform(static_cast<int>(m & 0xffffffffu)));
      if (m) return min(nq, word * 64 + __builtin_ctzll(m));
    }
    return nq;
  }

  // Receive terms of the 8 channels [rx_base, rx_base + 8), formed together:
  // lane l holds channel rx_base + l / 8 at tile position l % 8 (RC receives
  // along x, CR along y; a run may cross from RC to CR), and the min / max
  // keys over its quad of positions (0-3 or 4-7).  Live channels of a tile
  // are consecutive per aperture, so one fill serves up to 8 setups.
  int rx_base = -kBig;
  Term rx_term;
  Key rx_lo, rx_hi;

  __device__ void fill_receive(int q0) {
    const int lane = threadIdx.x & (kWave - 1);
    const int q = q0 + (lane >> 3);
    const int pos = lane & 7;
    Term r{kInactive, 0.f, 0.f, 0.f};
    if (q < 2 * p.nc) {
      const int cfg = q >= p.nc ? 1 : 0;
      const bool axis_x = cfg == 0;
      if (in_grid(axis_x, pos)) {
        const double e = static_cast<double>((cfg == 0 ? p.xe : p.ye)[q - cfg * p.nc]);
        r = receive_term(coord(axis_x, pos), zv, e, p.c, p.fs, p.fd, p.fn);
      }
    }
    const bool on = r.i != kInactive;
    Key lo = on ? Key{r.i, __float_as_uint(r.f)} : kKeyHigh;
    Key hi = on ? Key{r.i, __float_as_uint(r.f)} : kKeyLow;
    lo = key_min(lo, dpp_key<0xB1>(lo));  // quad_perm [1,0,3,2]
    hi = key_max(hi, dpp_key<0xB1>(hi));
    lo = key_min(lo, dpp_key<0x4E>(lo));  // quad_perm [2,3,0,1]
    hi = key_max(hi, dpp_key<0x4E>(hi));
    rx_term = r;
    rx_lo = lo;
    rx_hi = hi;
    rx_base = q0;
  }

  // Channel q as this wave sees it.  rx_pos: receive position (tile-local,
  // along x for RC, along y for CR) of this lane's voxel, given for RC and
  // CR; quarter g of the tile (x half g / 2, y half g % 2) for ch.g.  Called
  // by the whole wave with q uniform.
  __device__ Channel setup(int q, int rx_pos_rc, int rx_pos_cr, int g) {
    const int lane = threadIdx.x & (kWave - 1);
    Channel ch;
    ch.q = q;
    ch.cfg = q >= p.nc ? 1 : 0;
    if (q < rx_base || q >= rx_base + 8) fill_receive(q);
    const int slot = uniform(8 * (q - rx_base));  // first lane of channel q in the fill
    const int src = slot + (ch.cfg == 0 ? rx_pos_rc : rx_pos_cr);
    ch.rx.i = __shfl(rx_term.i, src, kWave);
    ch.rx.f = __shfl(rx_term.f, src, kWave);
    ch.rx.c = __shfl(rx_term.c, src, kWave);
    ch.rx.s = __shfl(rx_term.s, src, kWave);
    // Min / max of the receive keys: positions 0-3 and 4-7 (quads), then all 8.
    const Key hlo[2] = {Key{__builtin_amdgcn_readlane(rx_lo.i, slot),
                            static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(rx_lo.f), slot))},
                        Key{__builtin_amdgcn_readlane(rx_lo.i, slot + 4),
                            static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(rx_lo.f), slot + 4))}};
    const Key hhi[2] = {Key{__builtin_amdgcn_readlane(rx_hi.i, slot),
                            static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(rx_hi.f), slot))},
                        Key{__builtin_amdgcn_readlane(rx_hi.i, slot + 4),
                            static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(rx_hi.f), slot + 4))}};
    const Key tlo = key_min(hlo[0], hlo[1]), thi = key_max(hhi[0], hhi[1]);
    // Lane a: windows of angle a for the tile and for quarter g.
    const int a = lane < p.na ? lane : 0;
    const Key* tw = txw + (ch.cfg * kGridMaxAngles + a) * 6;
    int2 tile_w = window(tlo, thi, tw[0], tw[1], p.ns);
    // Quarter g: RC receive half g / 2 (x), transmit half g % 2 (y); CR the other way.
    const int rh = ch.cfg == 0 ? g >> 1 : g & 1;
    const int th = ch.cfg == 0 ? g & 1 : g >> 1;
    // (selects, not indexing: a runtime index would place the key arrays in scratch memory)
    int2 group_w = window(rh ? hlo[1] : hlo[0], rh ? hhi[1] : hhi[0], tw[2 + 2 * th], tw[3 + 2 * th], p.ns);
    if (lane >= p.na) tile_w = group_w = make_int2(kBig, -kBig);
    ch.w = Windows{tile_w.x, tile_w.y};
    ch.g = Windows{group_w.x, group_w.y};
    return ch;
  }

  // Tile windows of channel q (lane a: angle a; empty beyond na), its receive
  // terms filled as by setup() (lanes slot .. slot + 7, slot = 8 (q - rx_base)).
  // Called by the whole wave with q uniform.
  __device__ Windows tile_windows(int q) {
    const int lane = threadIdx.x & (kWave - 1);
    const int cfg = q >= p.nc ? 1 : 0;
    if (q < rx_base || q >= rx_base + 8) fill_receive(q);
    const int slot = uniform(8 * (q - rx_base));
    const Key lo0{__builtin_amdgcn_readlane(rx_lo.i, slot),
                  static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(rx_lo.f), slot))};
    const Key lo1{__builtin_amdgcn_readlane(rx_lo.i, slot + 4),
                  static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(rx_lo.f), slot + 4))};
    const Key hi0{__builtin_amdgcn_readlane(rx_hi.i, slot),
                  static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(rx_hi.f), slot))};
    const Key hi1{__builtin_amdgcn_readlane(rx_hi.i, slot + 4),
                  static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(rx_hi.f), slot + 4))};
    const int a = lane < p.na ? lane : 0;
    const Key* tw = txw + (cfg * kGridMaxAngles + a) * 6;
    int2 w = window(key_min(lo0, lo1), key_max(hi0, hi1), tw[0], tw[1], p.ns);
    if (lane >= p.na) w = make_int2(kBig, -kBig);
    return Windows{w.x, w.y};
  }

  // Transmit term of angle a at tile position pos (along y for RC, x for CR).
  __device__ Term transmit(int cfg, int a, int pos) const { return txt[(cfg * kGridMaxAngles + a) * 8 + pos]; }
};

}  // namespace rcb
