// Coefficient lists of the cmdata module (src/bsr/BSR_MAT3/mod_cmdata.f90):
// the block builder (Add_cdata, add_cdata.f90:2-56) and the merge of sorted
// blocks (Merge_CDATA, merge_cdata.f90:2-91), host side.  Keys are the
// tuples (K1, K2, K3, K4) compared lexicographically; indices are 1-based
// as in the reference.
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/bsr_gpu.h"

namespace bsrg {
int fail(int code, const std::string& msg);
}

namespace {
using Key = std::tuple<int, int, int, int>;
}

// n successive Add_cdata calls into an empty block: the records sorted by key
// (a stable sort, so equal keys keep their arrival order) and each key's
// coefficients summed in that order, which is the order in which Add_cdata
// adds them into the slot of the first arrival.
int bsrg_cdata_build(int n, const int* j1, const int* j2, const int* j3, const int* j4, const double* c, int* k1,
                     int* k2, int* k3, int* k4, double* cout, int* nout) {
  if (n < 0 || !nout) return bsrg::fail(-2, "bsrg_cdata_build: bad arguments");
  *nout = 0;
  if (n == 0) return 0;
  if (!j1 || !j2 || !j3 || !j4 || !c || !k1 || !k2 || !k3 || !k4 || !cout)
    return bsrg::fail(-2, "bsrg_cdata_build: null argument");
  std::vector<int> p(n);
  std::iota(p.begin(), p.end(), 0);
  auto key = [&](int i) { return Key(j1[i], j2[i], j3[i], j4[i]); };
  std::stable_sort(p.begin(), p.end(), [&](int a, int b) { return key(a) < key(b); });
  int m = -1;
  for (int q = 0; q < n; ++q) {
    const int i = p[q];
    if (m >= 0 && Key(k1[m], k2[m], k3[m], k4[m]) == key(i)) {
      cout[m] += c[i];
      continue;
    }
    ++m;
    k1[m] = j1[i]; k2[m] = j2[i]; k3[m] = j3[i]; k4[m] = j4[i];
    cout[m] = c[i];
  }
  *nout = m + 1;
  return 0;
}

// Merge_CDATA over nn sorted blocks [ip(m), jp(m)] of the arrays k1..k4,
// cdata.  The reference repeats one step: if two block heads hold the same
// key, the first such pair (ii < jj, smallest ii, then smallest jj) is
// merged -- cdata(head ii) += cdata(head jj), cdata(head jj) = 0, block jj
// advances -- otherwise the smallest head is appended to ipt when
// |cdata| > eps and its block advances.  The same steps are taken here, with
// the heads kept in a map from key to the blocks whose head holds it, so a
// step costs O(log nn) instead of a scan of all pairs; the sums, their
// order and ipt are those of the reference.
int bsrg_cdata_merge(int nn, int* ip, const int* jp, int n, const int* k1, const int* k2, const int* k3,
                     const int* k4, double* cdata, double eps_c, int* ipt, int* nc) {
  if (nn < 0 || n < 0 || !nc) return bsrg::fail(-2, "bsrg_cdata_merge: bad arguments");
  *nc = 0;
  if (nn == 0) return 0;
  if (!ip || !jp || !k1 || !k2 || !k3 || !k4 || !cdata || !ipt) return bsrg::fail(-2, "bsrg_cdata_merge: null argument");
  for (int m = 0; m < nn; ++m)
    if (jp[m] > 0 && ip[m] > 0 && (ip[m] > n || jp[m] > n))
      return bsrg::fail(-2, "bsrg_cdata_merge: block outside the n records");
  auto key = [&](int i) { return Key(k1[i - 1], k2[i - 1], k3[i - 1], k4[i - 1]); };
  std::map<Key, std::set<int>> heads;  // key -> blocks (0-based) whose head holds it
  std::set<std::pair<int, Key>> clash;  // (smallest block, key) of keys held by two or more heads
  auto insert = [&](int m) {
    const Key k = key(ip[m]);
    auto& s = heads[k];
    if (s.size() >= 2) clash.erase({*s.begin(), k});
    s.insert(m);
    if (s.size() >= 2) clash.insert({*s.begin(), k});
  };
  auto remove = [&](int m) {
    const Key k = key(ip[m]);
    auto it = heads.find(k);
    auto& s = it->second;
    if (s.size() >= 2) clash.erase({*s.begin(), k});
    s.erase(m);
    if (s.empty()) heads.erase(it);
    else if (s.size() >= 2) clash.insert({*s.begin(), k});
  };
  for (int m = 0; m < nn; ++m) {
    if (jp[m] <= 0) ip[m] = 0;
    if (ip[m] > 0 && ip[m] <= jp[m]) insert(m);
    else ip[m] = 0;
  }
  auto advance = [&](int m) {
    remove(m);
    if (++ip[m] <= jp[m]) insert(m);
    else ip[m] = 0;
  };
  int cnt = 0;
  while (!heads.empty()) {
    if (!clash.empty()) {
      const auto& s = heads[clash.begin()->second];
      auto it = s.begin();
      const int ii = *it, jj = *std::next(it);
      cdata[ip[ii] - 1] += cdata[ip[jj] - 1];
      cdata[ip[jj] - 1] = 0.0;
      advance(jj);
      continue;
    }
    const int mm = *heads.begin()->second.begin();
    const int i = ip[mm];
    if (std::abs(cdata[i - 1]) > eps_c) ipt[cnt++] = i;
    advance(mm);
  }
  *nc = cnt;
  return 0;
}
