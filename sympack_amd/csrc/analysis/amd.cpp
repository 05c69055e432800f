// Approximate minimum degree ordering, as Ordering::AMD computes it
// (src/Ordering.cpp:411-525 -> amdbar.F: Amestoy, Davis & Duff's AMD with
// approximate external degrees, element absorption, mass elimination and
// hash-based supervariable detection; no aggressive absorption).
//
// The output permutation must be identical to symPACK's, so the quotient
// graph is kept in one workspace array `iw` of length nnz + n exactly as
// Ordering.cpp sizes it (IWLEN), with the same list discipline: degree lists
// insert at the head, hash buckets share the `head` array with the degree
// lists (negative heads), and the workspace is compacted when the new element
// would overflow it.  All arrays are 1-based.
#include <algorithm>
#include <cstdlib>
#include <limits>

#include "analysis.hpp"

namespace spk {

namespace {

using I64 = int64_t;

struct AmdWork {
  I64 n;
  std::vector<I64> pe, len, nv, next, last, head, elen, degree, w, iw;
  I64 iwlen, pfree;

  // unlink variable i from the degree list it is on
  void unlink(I64 i) {
    const I64 ilast = last[i], inext = next[i];
    if (inext != 0) last[inext] = ilast;
    if (ilast != 0)
      next[ilast] = inext;
    else
      head[degree[i]] = inext;
  }
  void push_degree_list(I64 i, I64 deg) {
    const I64 inext = head[deg];
    if (inext != 0) last[inext] = i;
    next[i] = inext;
    head[deg] = i;
  }
  void reset_marks(I64& wflg) {
    for (I64 x = 1; x <= n; ++x)
      if (w[x] != 0) w[x] = 1;
    wflg = 2;
  }
  // Compact iw: every live list (pe > 0) is moved to the front, then the
  // partial new element iw[pme1 .. pfree-1] after them.
  void compact(I64& pme1) {
    for (I64 j = 1; j <= n; ++j) {
      const I64 pn = pe[j];
      if (pn > 0) {
        pe[j] = iw[pn];
        iw[pn] = -j;
      }
    }
    I64 pdst = 1, psrc = 1;
    const I64 pend = pme1 - 1;
    while (psrc <= pend) {
      const I64 j = -iw[psrc];
      ++psrc;
      if (j > 0) {
        iw[pdst] = pe[j];
        pe[j] = pdst;
        ++pdst;
        const I64 lenj = len[j];
        for (I64 k = 0; k <= lenj - 2; ++k) iw[pdst + k] = iw[psrc + k];
        pdst += lenj - 1;
        psrc += lenj - 1;
      }
    }
    const I64 p1 = pdst;
    for (I64 q = pme1; q <= pfree - 1; ++q) iw[pdst++] = iw[q];
    pme1 = p1;
    pfree = pdst;
  }

  void run() {
    const I64 iovflo = std::numeric_limits<I64>::max();
    I64 wflg = 2, mindeg = 1, nel = 0, dmax = 0;
    const I64 hmod = std::max<I64>(1, n - 1);
    const I64 wbig = iovflo - n;
    for (I64 i = 1; i <= n; ++i) {
      last[i] = head[i] = elen[i] = 0;
      nv[i] = w[i] = 1;
      degree[i] = len[i];
    }
    for (I64 i = 1; i <= n; ++i) {
      if (degree[i] > 0) {
        push_degree_list(i, degree[i]);
      } else {  // isolated: ordered first
        ++nel;
        elen[i] = -nel;
        pe[i] = 0;
        w[i] = 0;
      }
    }

    while (nel < n) {
      // ---- pivot of minimum approximate degree ----
      I64 deg = mindeg, me = 0;
      for (; deg <= n; ++deg) {
        me = head[deg];
        if (me > 0) break;
      }
      mindeg = deg;
      {
        const I64 inext = next[me];
        if (inext != 0) last[inext] = 0;
        head[deg] = inext;
      }
      const I64 elenme = elen[me];
      elen[me] = -(nel + 1);
      I64 nvpiv = nv[me];
      nel += nvpiv;
      nv[me] = -nvpiv;
      I64 degme = 0;

      // ---- new element Lme ----
      I64 pme1, pme2, newmem;
      if (elenme == 0) {
        // in place: me was a variable with no adjacent elements
        pme1 = pe[me];
        pme2 = pme1 - 1;
        for (I64 p = pme1; p <= pme1 + len[me] - 1; ++p) {
          const I64 i = iw[p];
          const I64 nvi = nv[i];
          if (nvi > 0) {
            degme += nvi;
            nv[i] = -nvi;
            iw[++pme2] = i;
            unlink(i);
          }
        }
        newmem = 0;
      } else {
        // union of the adjacent elements' lists and me's variables, built
        // at the end of iw
        I64 p = pe[me];
        pme1 = pfree;
        const I64 slenme = len[me] - elenme;
        for (I64 knt1 = 1; knt1 <= elenme + 1; ++knt1) {
          I64 e, pj, ln;
          if (knt1 > elenme) {
            e = me;
            pj = p;
            ln = slenme;
          } else {
            e = iw[p++];
            pj = pe[e];
            ln = len[e];
          }
          for (I64 knt2 = 1; knt2 <= ln; ++knt2) {
            const I64 i = iw[pj++];
            const I64 nvi = nv[i];
            if (nvi <= 0) continue;
            if (pfree > iwlen) {
              pe[me] = p;
              len[me] -= knt1;
              if (len[me] == 0) pe[me] = 0;
              pe[e] = pj;
              len[e] = ln - knt2;
              if (len[e] == 0) pe[e] = 0;
              compact(pme1);
              pj = pe[e];
              p = pe[me];
            }
            degme += nvi;
            nv[i] = -nvi;
            iw[pfree++] = i;
            unlink(i);
          }
          if (e != me) {  // element e is absorbed into me
            pe[e] = -me;
            w[e] = 0;
          }
        }
        pme2 = pfree - 1;
        newmem = pfree - pme1;
      }
      degree[me] = degme;
      pe[me] = pme1;
      len[me] = pme2 - pme1 + 1;
      if (wflg >= wbig) reset_marks(wflg);

      // ---- |Le \ Lme| for every element e adjacent to Lme ----
      for (I64 pme = pme1; pme <= pme2; ++pme) {
        const I64 i = iw[pme];
        const I64 eln = elen[i];
        if (eln <= 0) continue;
        const I64 nvi = -nv[i];
        const I64 wnvi = wflg - nvi;
        for (I64 p = pe[i]; p <= pe[i] + eln - 1; ++p) {
          const I64 e = iw[p];
          I64 we = w[e];
          if (we >= wflg)
            we -= nvi;
          else if (we != 0)
            we = degree[e] + wnvi;
          w[e] = we;
        }
      }

      // ---- approximate degrees, list pruning, mass elimination, hashes ----
      for (I64 pme = pme1; pme <= pme2; ++pme) {
        const I64 i = iw[pme];
        const I64 p1 = pe[i];
        const I64 p2 = p1 + elen[i] - 1;
        I64 pn = p1, hash = 0, d = 0;
        for (I64 p = p1; p <= p2; ++p) {
          const I64 e = iw[p];
          const I64 we = w[e];
          if (we != 0) {
            d += we - wflg;
            iw[pn++] = e;
            hash += e;
          }
        }
        elen[i] = pn - p1 + 1;
        const I64 p3 = pn;
        for (I64 p = p2 + 1; p <= p1 + len[i] - 1; ++p) {
          const I64 j = iw[p];
          const I64 nvj = nv[j];
          if (nvj > 0) {
            d += nvj;
            iw[pn++] = j;
            hash += j;
          }
        }
        if (elen[i] == 1 && p3 == pn) {
          // only adjacent to me: eliminated together with it
          pe[i] = -me;
          const I64 nvi = -nv[i];
          degme -= nvi;
          nvpiv += nvi;
          nel += nvi;
          nv[i] = 0;
          elen[i] = 0;
        } else {
          degree[i] = std::min(degree[i], d);
          // me becomes the first element of i's list
          iw[pn] = iw[p3];
          iw[p3] = iw[p1];
          iw[p1] = me;
          len[i] = pn - p1 + 1;
          hash = hash % hmod + 1;
          const I64 j = head[hash];
          if (j <= 0) {
            next[i] = -j;
            head[hash] = -i;
          } else {
            next[i] = last[j];
            last[j] = i;
          }
          last[i] = hash;
        }
      }
      degree[me] = degme;
      dmax = std::max(dmax, degme);
      wflg += dmax;
      if (wflg >= wbig) reset_marks(wflg);

      // ---- supervariable detection within each hash bucket ----
      for (I64 pme = pme1; pme <= pme2; ++pme) {
        I64 i = iw[pme];
        if (nv[i] >= 0) continue;
        const I64 hash = last[i];
        const I64 j0 = head[hash];
        if (j0 == 0) continue;
        if (j0 < 0) {
          i = -j0;
          head[hash] = 0;
        } else {
          i = last[j0];
          last[j0] = 0;
        }
        while (i != 0 && next[i] != 0) {
          const I64 ln = len[i], eln = elen[i];
          for (I64 p = pe[i] + 1; p <= pe[i] + ln - 1; ++p) w[iw[p]] = wflg;
          I64 jlast = i, j = next[i];
          while (j != 0) {
            bool same = len[j] == ln && elen[j] == eln;
            for (I64 p = pe[j] + 1; same && p <= pe[j] + ln - 1; ++p) same = w[iw[p]] == wflg;
            if (same) {  // j is indistinguishable from i: absorb it
              pe[j] = -i;
              nv[i] += nv[j];
              nv[j] = 0;
              elen[j] = 0;
              j = next[j];
              next[jlast] = j;
            } else {
              jlast = j;
              j = next[j];
            }
          }
          ++wflg;
          i = next[i];
        }
      }

      // ---- finalize degrees, put Lme's variables back on degree lists ----
      I64 p = pme1;
      const I64 nleft = n - nel;
      for (I64 pme = pme1; pme <= pme2; ++pme) {
        const I64 i = iw[pme];
        const I64 nvi = -nv[i];
        if (nvi <= 0) continue;
        nv[i] = nvi;
        const I64 d = std::max<I64>(1, std::min(degree[i] + degme - nvi, nleft - nvi));
        push_degree_list(i, d);
        last[i] = 0;
        mindeg = std::min(mindeg, d);
        degree[i] = d;
        iw[p++] = i;
      }
      nv[me] = nvpiv + degme;
      len[me] = p - pme1;
      if (len[me] == 0) {
        pe[me] = 0;
        w[me] = 0;
      }
      if (newmem != 0) pfree = p;
    }

    // ---- order of the absorbed variables from the assembly tree ----
    for (I64 i = 1; i <= n; ++i) {
      if (elen[i] != 0) continue;
      I64 j = -pe[i];
      while (elen[j] >= 0) j = -pe[j];
      const I64 e = j;
      I64 k = -elen[e];
      j = i;
      while (elen[j] >= 0) {
        const I64 jnext = -pe[j];
        pe[j] = -e;
        if (elen[j] == 0) elen[j] = k++;
        j = jnext;
      }
      elen[e] = -k;
    }
    for (I64 i = 1; i <= n; ++i) {
      const I64 k = std::llabs(elen[i]);
      last[k] = i;
      elen[i] = k;
    }
  }
};

}  // namespace

void order_amd(const Graph& g, Ordering& ord) {
  if (g.keep_diag)
    throw std::logic_error(
        "SparseMatrixGraph must be expanded and not including the diagonal in order to call AMD\n");
  const I64 n = g.n;
  ord.perm.assign(n, 0);
  ord.invp.assign(n, 0);
  if (n == 0) return;
  const I64 nnz = (I64)g.rowind.size();
  AmdWork a;
  a.n = n;
  a.iwlen = nnz + n;  // Ordering.cpp: IWLEN = colptr[N-1] + len(N) + N - 1 (1-based)
  a.pfree = nnz + 1;
  a.pe.assign(n + 1, 0);
  a.len.assign(n + 1, 0);
  for (I64 i = 1; i <= n; ++i) {
    a.pe[i] = g.colptr[i - 1];
    a.len[i] = g.colptr[i] - g.colptr[i - 1];
  }
  a.iw.assign(a.iwlen + 2, 0);
  for (I64 p = 1; p <= nnz; ++p) a.iw[p] = g.rowind[p - 1];
  a.nv.assign(n + 1, 0);
  a.next.assign(n + 1, 0);
  a.last.assign(n + 1, 0);
  a.head.assign(n + 2, 0);
  a.elen.assign(n + 1, 0);
  a.degree.assign(n + 1, 0);
  a.w.assign(n + 1, 0);
  a.run();
  // invp = ELEN (1-based position of each vertex), perm rebuilt from it
  // (Ordering.cpp:519-522)
  for (I64 i = 1; i <= n; ++i) ord.invp[i - 1] = (Int)a.elen[i];
  for (I64 i = 1; i <= n; ++i) ord.perm[ord.invp[i - 1] - 1] = (Int)i;
}

}  // namespace spk
