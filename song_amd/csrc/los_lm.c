/*
 * los_lm.c — (l,m) offset helpers.
 *
 * SONG stores multipole hierarchies as flat arrays ordered l-major with m
 * ascending, keeping only m taken from the requested m-list with m <= l
 * (include/transfer2.h:141-147, include/perturbations2.h:611-644). The helpers
 * that compute these offsets live in the un-vendored CLASS fork (song_tools.c:
 * size_l_indexm, multipole2offset_l_indexm, size_indexl_indexm,
 * multipole2offset_indexl_indexm, offset2multipole_indexl_indexm, called at
 * source/transfer2.c:900, :1154, :1207 and source/perturbations2.c:1159);
 * these are restatements of that documented ordering.
 */
#include "los_internal.h"

/* Number of (l,m) with 0 <= l <= l_max and m in m_vec, m <= l. */
int los_size_l_indexm(int l_max, const int *m_vec, int m_size)
{
  int count = 0;
  for (int l = 0; l <= l_max; ++l)
    for (int im = 0; im < m_size; ++im)
      if (m_vec[im] <= l) ++count;
  return count;
}

/* Offset of (l,m) in the l = 0,1,2,... ordering; -1 if (l,m) is not stored. */
int los_offset_l_indexm(int l, int m, const int *m_vec, int m_size)
{
  int offset = 0;
  for (int ll = 0; ll <= l; ++ll)
    for (int im = 0; im < m_size; ++im) {
      if (m_vec[im] > ll) continue;
      if (ll == l && m_vec[im] == m) return offset;
      ++offset;
    }
  return -1;
}

/* Number of (l,m) with l in l_vec and m in m_vec, m <= l. */
int los_size_indexl_indexm(const int *l_vec, int l_size, const int *m_vec, int m_size)
{
  int count = 0;
  for (int il = 0; il < l_size; ++il)
    for (int im = 0; im < m_size; ++im)
      if (m_vec[im] <= l_vec[il]) ++count;
  return count;
}

/* Offset of (l,m) in the l_vec-major ordering; -1 if not stored. */
int los_offset_indexl_indexm(int l, int m, const int *l_vec, int l_size, const int *m_vec,
                             int m_size)
{
  int offset = 0;
  for (int il = 0; il < l_size; ++il)
    for (int im = 0; im < m_size; ++im) {
      if (m_vec[im] > l_vec[il]) continue;
      if (l_vec[il] == l && m_vec[im] == m) return offset;
      ++offset;
    }
  return -1;
}

/* Inverse of los_offset_indexl_indexm: indices (into l_vec, m_vec) of an offset. */
int los_offset2multipole_indexl_indexm(int offset, const int *l_vec, int l_size,
                                       const int *m_vec, int m_size, int *index_l,
                                       int *index_m)
{
  int o = 0;
  for (int il = 0; il < l_size; ++il)
    for (int im = 0; im < m_size; ++im) {
      if (m_vec[im] > l_vec[il]) continue;
      if (o == offset) {
        *index_l = il;
        *index_m = im;
        return _SUCCESS_;
      }
      ++o;
    }
  *index_l = -1;
  *index_m = -1;
  return _FAILURE_;
}
