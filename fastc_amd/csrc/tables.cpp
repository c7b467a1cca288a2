// Host-built lookup tables for the BC7 kernels.
//
// The reference ships two of these tables as literal arrays together with the
// programs that generated them (BPTCEncoder/src/BCLookupTables.h:26-89 for
// Optimal7CompressBC7Mode5, :365-446 for Optimal6CompressDXT1).  We rebuild
// them from those generator programs' definitions.  The third table caches
// CompressSingleColor's exhaustive endpoint search
// (BPTCEncoder/src/Compressor.cpp:268-331) for every channel value.
#include <stdint.h>
#include <string.h>

#include "internal.h"

namespace fh {

namespace {

// Distinct 8-bit values reachable by an nbits endpoint with bit replication.
int expanded_values(int nbits, uint32_t *vals) {
  uint32_t last = 0xFFFFFFFFu;
  int n = 0;
  for (uint32_t i = 0; i < 256; i++) {
    uint32_t num = (i >> (8 - nbits)) << (8 - nbits);
    num |= i >> nbits;
    if (num != last) {
      last = num;
      vals[n++] = num;
    }
  }
  return n;
}

// Weight of the second endpoint at index 1 for 2/3/4-bit indices
// (BPTCEncoder/src/Compressor.cpp:159-168); the first endpoint gets 64 - w.
const uint32_t kIndexOneWeight[4] = {31, 21, 9, 4};

}  // namespace

void build_bc7_mode5_table(uint8_t out[256][2]) {
  uint32_t v[256];
  const int nv = expanded_values(7, v);
  for (uint32_t i = 0; i < 256; i++) {
    uint32_t best = 0xFFFFFFFFu, bj = 0, bk = 0;
    for (int j = 0; j < nv; j++)
      for (int k = 0; k < nv; k++) {
        const uint32_t c = (43 * v[j] + 21 * v[k] + 32) >> 6;
        const uint32_t d = i > c ? i - c : c - i;
        if (d < best) {
          best = d;
          bj = j;
          bk = k;
        }
      }
    out[i][0] = static_cast<uint8_t>(v[bj] >> 1);
    out[i][1] = static_cast<uint8_t>(v[bk] >> 1);
  }
}

void build_bc7_dxt6_table(uint8_t out[256][2][3]) {
  uint32_t v[256];
  const int nv = expanded_values(6, v);
  for (uint32_t i = 0; i < 256; i++) {
    // entry 0: midpoint (a+b)/2 flagged 1; entry 1: one third (2a+b)/3 flagged 0
    uint32_t flag[2] = {1, 0}, lo[2], hi[2], dist[2];
    for (int w = 0; w < 2; w++) {
      uint32_t best = 0xFFFFFFFFu, bj = 0, bk = 0;
      for (int j = 0; j < nv; j++)
        for (int k = j; k < nv; k++) {
          const uint32_t c = w == 0 ? (v[j] + v[k]) / 2 : (2 * v[j] + v[k]) / 3;
          const uint32_t d = i > c ? i - c : c - i;
          if (d < best) {
            best = d;
            bj = j;
            bk = k;
          }
        }
      lo[w] = v[bj];
      hi[w] = v[bk];
      dist[w] = best;
    }
    // The closer of the two comes first; ties put the one-third entry first.
    const int first = dist[1] > dist[0] ? 0 : 1;
    for (int t = 0; t < 2; t++) {
      const int w = t == 0 ? first : 1 - first;
      out[i][t][0] = static_cast<uint8_t>(flag[w]);
      out[i][t][1] = static_cast<uint8_t>(lo[w] >> 2);
      out[i][t][2] = static_cast<uint8_t>(hi[w] >> 2);
    }
  }
}

// Entry index: ((nbits - 4) * 5 + cfg) * 4 + bpi, then * 256 + value, where
// cfg 0 = no p-bit, 1 + 2*pl + ph = p-bits (pl for the low endpoint, ph for
// the high one) and bpi = index bits - 1.  Entry = lo | hi << 8 | dist << 16,
// the first (lo, hi) pair in the reference's scan order that minimises
// |interp(lo, hi) - value| at index 1.
void build_bc7_single_color_table(uint32_t *out) {
  memset(out, 0, sizeof(uint32_t) * kSingleColorEntries);
  for (int nbits0 = 4; nbits0 <= 8; nbits0++)
    for (int cfg = 0; cfg < 5; cfg++)
      for (int bpi = 1; bpi < 4; bpi++) {
        const int have_pbit = cfg != 0;
        const int pl = have_pbit ? ((cfg - 1) >> 1) : 0;
        const int ph = have_pbit ? ((cfg - 1) & 1) : 0;
        const int nposs = 1 << nbits0;
        const int nbits = nbits0 + have_pbit;
        int vh[256], vl[256];
        for (int i = 0; i < nposs; i++) {
          int h = i, l = i;
          if (have_pbit) {
            h = (h << 1) | ph;
            l = (l << 1) | pl;
          }
          vh[i] = h << (8 - nbits);
          vh[i] |= vh[i] >> nbits;
          vl[i] = l << (8 - nbits);
          vl[i] |= vl[i] >> nbits;
        }
        const uint32_t w1 = kIndexOneWeight[bpi], w0 = 64 - w1;
        for (uint32_t val = 0; val < 256; val++) {
          uint32_t bestd = 0xFF, bi = 0xFF, bj = 0xFF;
          for (int i = 0; bestd > 0 && i < nposs; i++)
            for (int j = 0; bestd > 0 && j < nposs; j++) {
              const uint32_t c = (w0 * (uint32_t)vl[i] + w1 * (uint32_t)vh[j] + 32) >> 6;
              const uint32_t e = c > val ? c - val : val - c;
              if (e < bestd) {
                bestd = e;
                bi = (uint32_t)vl[i];
                bj = (uint32_t)vh[j];
              }
            }
          const int idx = (((nbits0 - 4) * 5 + cfg) * 4 + bpi) * 256 + (int)val;
          out[idx] = bi | (bj << 8) | (bestd << 16);
        }
      }
}

}  // namespace fh
