// Reverse Cuthill-McKee element ordering, host side of the structured-mesh init.
//
// Restates the George & Liu GENRCM algorithm (A. George, J. W. Liu, "Computer
// Solution of Large Sparse Positive Definite Systems", Prentice Hall 1981) with
// the exact tie-breaking of the reference's version, SRC/rcm.f90:
//   genrcm    SRC/rcm.f90:155-250   one RCM ordering per connected component
//   root_find SRC/rcm.f90:683-869   pseudo-peripheral root (last level, min degree)
//   level_set SRC/rcm.f90:333-466   breadth-first level structure
//   degree    SRC/rcm.f90:14-152    degrees inside the component
//   rcm       SRC/rcm.f90:507-681   Cuthill-McKee with an insertion sort of each
//                                   node's new neighbours by degree; the sort never
//                                   moves the first new neighbour (loop condition
//                                   "fnbr < l"), then the order is reversed.
// The element permutation feeds the global GLL numbering, which must be bit-exact
// with the reference (SRC/mesh_structured.f90:204-269 builds the adjacency).
//
// Arrays follow the reference convention: 1-based adjacency (adj_row[n+1],
// adj[nadj]) and a 1-based permutation perm[k] = old index of new element k.
#include <vector>
#include <algorithm>

namespace {

struct Graph {
  int n;
  const int* row;  // 1-based CSR, row[i-1] .. row[i]-1 (1-based positions)
  const int* adj;
  int begin(int node) const { return row[node - 1]; }
  int end(int node) const { return row[node]; }  // exclusive
  int nbr(int pos) const { return adj[pos - 1]; }
};

// level structure from root over masked nodes; returns number of levels.
// level[0..iccsze-1] holds the nodes, level_row[l] the first index of level l.
int level_set(const Graph& g, int root, std::vector<int>& mask, std::vector<int>& level_row,
              int* level, int& iccsze) {
  mask[root] = 0;
  level[0] = root;
  int level_num = 0, lvlend = 0;
  iccsze = 1;
  for (;;) {
    int lbegin = lvlend + 1;
    lvlend = iccsze;
    ++level_num;
    level_row[level_num] = lbegin;
    for (int i = lbegin; i <= lvlend; ++i) {
      int node = level[i - 1];
      for (int p = g.begin(node); p < g.end(node); ++p) {
        int nb = g.nbr(p);
        if (mask[nb] != 0) {
          ++iccsze;
          level[iccsze - 1] = nb;
          mask[nb] = 0;
        }
      }
    }
    if (iccsze - lvlend <= 0) break;
  }
  level_row[level_num + 1] = lvlend + 1;
  for (int i = 0; i < iccsze; ++i) mask[level[i]] = 1;
  return level_num;
}

void root_find(const Graph& g, int& root, std::vector<int>& mask, std::vector<int>& level_row,
               int* level) {
  int iccsze = 0;
  int level_num = level_set(g, root, mask, level_row, level, iccsze);
  if (level_num == 1 || level_num == iccsze) return;
  for (;;) {
    int mindeg = iccsze;
    int jstrt = level_row[level_num];
    root = level[jstrt - 1];
    if (jstrt < iccsze) {
      for (int j = jstrt; j <= iccsze; ++j) {
        int node = level[j - 1];
        int ndeg = 0;
        for (int p = g.begin(node); p < g.end(node); ++p)
          if (mask[g.nbr(p)] > 0) ++ndeg;
        if (ndeg < mindeg) {
          root = node;
          mindeg = ndeg;
        }
      }
    }
    int icc2 = 0;
    int level_num2 = level_set(g, root, mask, level_row, level, icc2);
    if (level_num2 <= level_num) break;
    level_num = level_num2;
    if (iccsze <= level_num) break;
  }
}

// degrees of the component of root (masked nodes only); fills ls by levels.
void degree(const Graph& g, int root, const std::vector<int>& mask, std::vector<int>& deg,
            std::vector<char>& seen, int* ls, int& iccsze) {
  ls[0] = root;
  seen[root] = 1;
  int lvlend = 0;
  iccsze = 1;
  for (;;) {
    int lbegin = lvlend + 1;
    lvlend = iccsze;
    for (int i = lbegin; i <= lvlend; ++i) {
      int node = ls[i - 1];
      int ideg = 0;
      for (int p = g.begin(node); p < g.end(node); ++p) {
        int nb = g.nbr(p);
        if (mask[nb] != 0) {
          ++ideg;
          if (!seen[nb]) {
            seen[nb] = 1;
            ++iccsze;
            ls[iccsze - 1] = nb;
          }
        }
      }
      deg[node] = ideg;
    }
    if (iccsze - lvlend == 0) break;
  }
  for (int i = 0; i < iccsze; ++i) seen[ls[i]] = 0;
}

void rcm(const Graph& g, int root, std::vector<int>& mask, std::vector<int>& deg,
         std::vector<char>& seen, int* perm, int& iccsze) {
  degree(g, root, mask, deg, seen, perm, iccsze);
  mask[root] = 0;
  if (iccsze <= 1) return;
  int lvlend = 0, lnbr = 1;
  while (lvlend < lnbr) {
    int lbegin = lvlend + 1;
    lvlend = lnbr;
    for (int i = lbegin; i <= lvlend; ++i) {
      int node = perm[i - 1];
      int fnbr = lnbr + 1;
      for (int p = g.begin(node); p < g.end(node); ++p) {
        int nb = g.nbr(p);
        if (mask[nb] != 0) {
          ++lnbr;
          mask[nb] = 0;
          perm[lnbr - 1] = nb;
        }
      }
      if (lnbr <= fnbr) continue;
      int k = fnbr;
      while (k < lnbr) {
        int l = k;
        ++k;
        int nb = perm[k - 1];
        while (fnbr < l) {
          int lperm = perm[l - 1];
          if (deg[lperm] <= deg[nb]) break;
          perm[l] = lperm;
          --l;
        }
        perm[l] = nb;
      }
    }
  }
  std::reverse(perm, perm + iccsze);
}

}  // namespace

extern "C" int sem2d_genrcm(int node_num, int adj_num, const int* adj_row, const int* adj,
                            int* perm) {
  if (node_num <= 0 || !adj_row || !perm || (adj_num > 0 && !adj)) return 1;
  if (adj_row[node_num] - 1 != adj_num) return 2;
  Graph g{node_num, adj_row, adj};
  std::vector<int> mask(node_num + 1, 1), deg(node_num + 1, 0), level_row(node_num + 2, 0);
  std::vector<char> seen(node_num + 1, 0);
  mask[0] = 0;
  int num = 1;
  for (int i = 1; i <= node_num; ++i) {
    if (mask[i] == 0) continue;
    int root = i;
    root_find(g, root, mask, level_row, perm + (num - 1));
    int iccsze = 0;
    rcm(g, root, mask, deg, seen, perm + (num - 1), iccsze);
    num += iccsze;
    if (node_num < num) return 0;
  }
  return 0;
}
