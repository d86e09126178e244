/*
 * rts_remap.h -- value encoding of columns (the reference's remap surface, remap.cpp:110-291,
 * declared at bindex.h:148-152 and rt.h:136-140).
 *
 * C ABI: GPU encoders over device columns plus the two map lookups, on plain arrays.  The
 * encode map of a column is a pair of arrays: its distinct values in increasing order and
 * the start code of each.
 *
 * C++ callers additionally get the reference's names as inline functions over the C ABI,
 * with the reference's global `encodeMap` (remap.cpp:3).  Header quirks resolved here:
 * encodeQuery takes the 3-argument form of remap.cpp:243 / rt.h:138 (bindex.h:150 declares
 * a 2-argument one that nothing defines), and findKeyByValue takes std::map<CODE, CODE>
 * as remap.cpp:274 defines it (bindex.h:152 and rt.h:140 declare std::map<CODE, int>).
 */
#ifndef RTS_REMAP_H_
#define RTS_REMAP_H_

#include "rtscan_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Rank form of normalEncode (remap.cpp:110-203) for one device column of n rows.
 * The row at sorted index i gets code encode_min + i * I, I = (int)((encode_max -
 * encode_min) / n) computed in double, 32-bit wrap-around; rows of equal value keep their
 * row order (the reference's sort leaves it unspecified, bindex.h:154-162).
 *   data          device, n values (read)
 *   codes         device, n codes in row order (may be `data`: encodes in place, as the
 *                 reference overwrites initialDataSet[column], remap.cpp:186-197)
 *   sorted_pos    device, n row indices in sorted order, or NULL (the reference's
 *                 sorted_pos[column])
 *   sorted_codes  device, n codes in sorted order, or NULL (sorted_data[column])
 *   map_values / map_starts   device, room for n entries; the first *map_size receive the
 *                 distinct values and their start codes (encodeMap[column])
 *   map_size      host; written after the distinct values are counted (the call
 *                 synchronises the stream once for it; the last kernel stays queued).
 * n must be in [1, 2^31 - 1]. */
hipError_t rts_encode_column(const CODE *data, int64_t n, CODE encode_min, CODE encode_max, CODE *codes,
                             POSTYPE *sorted_pos, CODE *sorted_codes, CODE *map_values,
                             CODE *map_starts, int64_t *map_size, hipStream_t stream);

/* Dispersed form of normalEncode (remap.cpp:205-241) for one device column.  Value v with
 * c rows owns [start(v), start(v) + step(v) - 1), step(v) = (int)(D * c), D = (encode_max -
 * encode_min) / n in double, start(v) = encode_min + sum of step(u) over u < v (mod 2^32);
 * each row's offset in that range is hash(seed, row) % (step(v) - 1), or 0 when step(v) < 2
 * (the reference draws rand() and divides by zero when step(v) == 1).  Arguments as above;
 * returns after the work has finished. */
hipError_t rts_encode_column_dispersed(const CODE *data, int64_t n, CODE encode_min, CODE encode_max,
                                       uint64_t seed, CODE *codes, CODE *map_values, CODE *map_starts,
                                       int64_t *map_size, hipStream_t stream);

/* encodeQuery (remap.cpp:243-272) over a host map of m entries: LT and GE give the start
 * code of the first value >= old_query, GT that of the first value > old_query, LE that
 * start minus one; `sentinel` when there is no such value (the reference returns its row
 * count N).  EQ and BT return hipErrorInvalidValue (the reference prints and exits). */
hipError_t rts_encode_lookup(const CODE *map_values, const CODE *map_starts, int64_t m, CODE old_query,
                             enum OPERATOR op, CODE sentinel, CODE *code);

/* findKeyByValue (remap.cpp:274-283) over a host map: the value of the last entry whose
 * start code is <= code when a later entry starts above code; 0 otherwise. */
CODE rts_encode_key_of(const CODE *map_values, const CODE *map_starts, int64_t m, CODE code);

#ifdef __cplusplus
}
#endif

/* The library itself defines no C++ symbols: its sources set RTS_REMAP_NO_CXX. */
#if defined(__cplusplus) && !defined(RTS_REMAP_NO_CXX)
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#ifndef MAX_BINDEX_NUM
#define MAX_BINDEX_NUM 256 /* bindex.h:33 */
#endif

/* remap.cpp:3 -- distinct value -> start code, per column. */
inline std::map<CODE, CODE> encodeMap[MAX_BINDEX_NUM];
/* Row count of each column's last encode: encodeQuery's "no such value" answer (the
 * reference returns its compile-time N, which its callers pass as data_num,
 * rtscan.cpp:1072, rtc3.cpp:934-936). */
inline CODE encodeRows[MAX_BINDEX_NUM];

namespace rts_remap {

struct DeviceArray {
  void *p = nullptr;
  explicit DeviceArray(size_t bytes) {
    if (rts_malloc(&p, bytes) != hipSuccess) p = nullptr;
  }
  ~DeviceArray() {
    if (p) (void)rts_free(p);
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

/* Encode host column `data` in place on the GPU; optionally return the sorted positions and
 * codes in malloc'd arrays; refresh encodeMap[column_id]. */
inline bool encode_host_column(CODE *data, int column_id, CODE encode_min, CODE encode_max, int n,
                               bool dispersed, POSTYPE **sorted_pos, CODE **sorted_data) {
  if (n <= 0 || column_id < 0 || column_id >= MAX_BINDEX_NUM) return false;
  const size_t bytes = (size_t)n * sizeof(CODE);
  DeviceArray col(bytes), pos(dispersed ? 1 : bytes), sc(dispersed ? 1 : bytes), mv(bytes), ms(bytes);
  if (!col.p || !pos.p || !sc.p || !mv.p || !ms.p) return false;
  int64_t m = 0;
  if (rts_memcpy_htod(col.p, data, bytes) != hipSuccess) return false;
  hipError_t e = dispersed
      ? rts_encode_column_dispersed(col.as<CODE>(), n, encode_min, encode_max, (uint64_t)column_id,
                                    col.as<CODE>(), mv.as<CODE>(), ms.as<CODE>(), &m, 0)
      : rts_encode_column(col.as<CODE>(), n, encode_min, encode_max, col.as<CODE>(), pos.as<POSTYPE>(),
                          sc.as<CODE>(), mv.as<CODE>(), ms.as<CODE>(), &m, 0);
  if (e != hipSuccess || rts_stream_synchronize(0) != hipSuccess) return false;
  std::vector<CODE> values((size_t)m), starts((size_t)m);
  if (rts_memcpy_dtoh(data, col.p, bytes) != hipSuccess ||
      rts_memcpy_dtoh(values.data(), mv.p, (size_t)m * sizeof(CODE)) != hipSuccess ||
      rts_memcpy_dtoh(starts.data(), ms.p, (size_t)m * sizeof(CODE)) != hipSuccess)
    return false;
  if (!dispersed) {
    POSTYPE *hp = (POSTYPE *)malloc(bytes);
    CODE *hs = (CODE *)malloc(bytes);
    if (!hp || !hs || rts_memcpy_dtoh(hp, pos.p, bytes) != hipSuccess ||
        rts_memcpy_dtoh(hs, sc.p, bytes) != hipSuccess) {
      free(hp);
      free(hs);
      return false;
    }
    sorted_pos[column_id] = hp;
    sorted_data[column_id] = hs;
  }
  std::map<CODE, CODE> &em = encodeMap[column_id];
  em.clear();
  for (int64_t i = 0; i < m; i++) em.emplace_hint(em.end(), values[(size_t)i], starts[(size_t)i]);
  encodeRows[column_id] = (CODE)n;
  return true;
}

}  // namespace rts_remap

/* normalEncode, rank form (remap.cpp:110-203): encodes initialDataSet[0..column_num) in
 * place, sets sorted_pos[c] / sorted_data[c] to malloc'd arrays of data_num entries and
 * fills encodeMap[c].  Returns initialDataSet, or NULL when a device call fails. */
inline CODE **normalEncode(CODE **initialDataSet, int column_num, CODE encode_min, CODE encode_max,
                           int data_num, POSTYPE **sorted_pos, CODE **sorted_data) {
  for (int c = 0; c < column_num; c++)
    if (!rts_remap::encode_host_column(initialDataSet[c], c, encode_min, encode_max, data_num, false,
                                       sorted_pos, sorted_data))
      return NULL;
  return initialDataSet;
}

/* normalEncode, dispersed form (remap.cpp:205-241).  The codes are written into the
 * caller's arrays (the reference swaps in new malloc'd arrays and frees the old ones). */
inline CODE **normalEncode(CODE **initialDataSet, int column_num, CODE encode_min, CODE encode_max,
                           int data_num) {
  for (int c = 0; c < column_num; c++)
    if (!rts_remap::encode_host_column(initialDataSet[c], c, encode_min, encode_max, data_num, true,
                                       NULL, NULL))
      return NULL;
  return initialDataSet;
}

/* encodeQuery (remap.cpp:243-272).  "eq" and "bt" print the reference's message and exit,
 * as the reference does. */
inline CODE encodeQuery(int column_id, CODE old_query, std::string &cmd) {
  const std::map<CODE, CODE> &em = encodeMap[column_id];
  std::map<CODE, CODE>::const_iterator it;
  if (cmd == "lt" || cmd == "ge") {
    it = em.lower_bound(old_query);
    return it == em.end() ? encodeRows[column_id] : it->second;
  }
  if (cmd == "le" || cmd == "gt") {
    it = em.upper_bound(old_query);
    if (it == em.end()) return encodeRows[column_id];
    return cmd == "le" ? it->second - 1 : it->second;
  }
  printf("`eq` and `bt` are not directly supported now, and they can be transformed into `lt` + `gt`.\n");
  exit(-1);
}

/* findKeyByValue (remap.cpp:274-283). */
inline CODE findKeyByValue(const CODE Val, std::map<CODE, CODE> &map_) {
  CODE last = 0;
  for (std::map<CODE, CODE>::const_iterator it = map_.begin(); it != map_.end(); ++it) {
    if (it->second > Val) return last;
    last = it->first;
  }
  return 0;
}

/* ifEncodeEqual (remap.cpp:285-291): both codes map to the same key. */
inline bool ifEncodeEqual(const CODE val1, const CODE val2, int bindex_id) {
  return findKeyByValue(val1, encodeMap[bindex_id]) == findKeyByValue(val2, encodeMap[bindex_id]);
}

#endif /* __cplusplus && !RTS_REMAP_NO_CXX */
#endif /* RTS_REMAP_H_ */
