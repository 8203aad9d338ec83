// h3dmg_layout.cpp -- the multi-block layout files of Hydro3D on the host:
// infodom.cin (block coordinates), mdmap.cin (blocks per rank) and the
// neighbour search over them, producing h3dmg_block tables with shared faces.
//
// Reference: read_mdmap and read_infodom, src_v70/alloc_dom.for:284-1458;
// block cell counts and spacing, src_v70/localparameters.for:80-104,247-249;
// boundary codes per block, src_v70/initial.for:98-100 (control.cin).
//
// Fortran list-directed input is read as the reference reads it: each READ
// starts on a new record and takes as many values as it needs, continuing
// on the following records when one runs out; whatever is left of the last
// record it touched is skipped.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/h3dmg.h"

namespace {

thread_local std::string g_layout_error;

h3dmg_status lfail(h3dmg_status st, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
h3dmg_status lfail(h3dmg_status st, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_layout_error = buf;
  return st;
}

// records of a text file, read statement by statement
struct Records {
  std::vector<std::string> lines;
  size_t next = 0;

  bool load(const char *path) {
    std::ifstream in(path);
    if (!in) return false;
    std::string l;
    while (std::getline(in, l)) lines.push_back(l);
    return true;
  }
  // the tokens of one list-directed READ of `count` values (separators:
  // blanks, tabs and commas); false at the end of the file
  bool read(size_t count, std::vector<std::string> &out) {
    out.clear();
    while (out.size() < count) {
      if (next >= lines.size()) return false;
      std::string l = lines[next++];
      for (char &c : l)
        if (c == ',' || c == '\t' || c == '\r') c = ' ';
      std::istringstream ss(l);
      std::string t;
      while (out.size() < count && ss >> t) out.push_back(t);
    }
    return true;
  }
  // skip one record (a READ into a dummy line or a bare READ)
  bool skip() {
    if (next >= lines.size()) return false;
    ++next;
    return true;
  }
};

bool to_int(const std::string &t, int &v) {
  char *e = nullptr;
  const long x = std::strtol(t.c_str(), &e, 10);
  if (e == t.c_str() || *e) return false;
  v = (int)x;
  return true;
}

bool to_double(const std::string &t, double &v) {
  std::string u = t;
  for (char &c : u)
    if (c == 'd' || c == 'D') c = 'e';  // Fortran exponent letters
  char *e = nullptr;
  v = std::strtod(u.c_str(), &e);
  return e != u.c_str() && !*e;
}

struct Domain {
  int rdiv = 1;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  int rank = -1;
};

}  // namespace

struct h3dmg_layout {
  std::vector<Domain> dom;
  int nranks = 0;
  std::vector<std::vector<int>> by_rank;  // domains of each rank, in mdmap order
};

namespace {

// read_mdmap, alloc_dom.for:297-374: type 1 lists "rank count ids..." per
// processor; type 2 gives ndoms/nprocs consecutive domains per rank and the
// remainder to the last one
h3dmg_status read_mdmap(const char *path, h3dmg_layout &L) {
  Records R;
  if (!R.load(path)) return lfail(H3DMG_ERR_ARG, "cannot open %s", path);
  std::vector<std::string> t;
  int type = 0, ndoms = 0, nprocs = 0;
  if (!R.read(1, t) || !to_int(t[0], type)) return lfail(H3DMG_ERR_ARG, "%s: mdmap type", path);
  if (!R.read(1, t) || !to_int(t[0], ndoms) || ndoms < 1)
    return lfail(H3DMG_ERR_ARG, "%s: number of domains", path);
  if (!R.read(1, t) || !to_int(t[0], nprocs) || nprocs < 1)
    return lfail(H3DMG_ERR_ARG, "%s: number of processors", path);
  if ((int)L.dom.size() != ndoms)
    return lfail(H3DMG_ERR_ARG, "%s: %d domains, infodom.cin has %zu", path, ndoms,
                 L.dom.size());
  L.nranks = nprocs;
  L.by_rank.assign(nprocs, {});
  if (type == 1) {
    R.skip();
    for (int p = 0; p < nprocs; ++p) {
      int rank = 0, nb = 0;
      if (!R.read(2, t) || !to_int(t[0], rank) || !to_int(t[1], nb) || rank < 0 ||
          rank >= nprocs || nb < 0)
        return lfail(H3DMG_ERR_ARG, "%s: processor line %d", path, p + 1);
      // the same READ continues with the nb domain ids
      --R.next;
      if (!R.read(2 + (size_t)nb, t)) return lfail(H3DMG_ERR_ARG, "%s: domain list", path);
      for (int q = 0; q < nb; ++q) {
        int id = 0;
        if (!to_int(t[2 + q], id) || id < 0 || id >= ndoms)
          return lfail(H3DMG_ERR_ARG, "%s: domain id %s", path, t[2 + q].c_str());
        L.dom[id].rank = rank;
        L.by_rank[rank].push_back(id);
      }
    }
  } else {
    const int per = ndoms / nprocs;
    int rem = ndoms - per * nprocs;
    if (ndoms % nprocs == 0) rem = per;
    for (int r = 0; r + 1 < nprocs; ++r)
      for (int i = 0; i < per; ++i) {
        L.dom[r * per + i].rank = r;
        L.by_rank[r].push_back(r * per + i);
      }
    for (int i = 0; i < rem; ++i) {
      const int id = (nprocs - 1) * per + i;
      if (id >= ndoms) break;
      L.dom[id].rank = nprocs - 1;
      L.by_rank[nprocs - 1].push_back(id);
    }
  }
  for (int i = 0; i < ndoms; ++i)
    if (L.dom[i].rank < 0)
      return lfail(H3DMG_ERR_ARG, "%s: domain %d on no processor (alloc_dom.for:398-405)", path, i);
  return H3DMG_OK;
}

// read_infodom, alloc_dom.for:424-481: type 1 lists "id rdiv x0 x1 y0 y1 z0
// z1" per domain after nine header records; type 2 splits Lx, Ly, Lz into
// idom x jdom x kdom equal domains, x fastest
h3dmg_status read_infodom(const char *path, h3dmg_layout &L) {
  Records R;
  if (!R.load(path)) return lfail(H3DMG_ERR_ARG, "cannot open %s", path);
  std::vector<std::string> t;
  int type = 0, ndoms = 0;
  if (!R.read(1, t) || !to_int(t[0], type)) return lfail(H3DMG_ERR_ARG, "%s: infodom type", path);
  if (!R.read(1, t) || !to_int(t[0], ndoms) || ndoms < 1)
    return lfail(H3DMG_ERR_ARG, "%s: number of domains", path);
  R.skip();
  L.dom.assign(ndoms, Domain{});
  if (type == 1) {
    for (int q = 0; q < 3; ++q) R.skip();
    int nd[3];
    for (int a = 0; a < 3; ++a)
      if (!R.read(1, t) || !to_int(t[0], nd[a]))
        return lfail(H3DMG_ERR_ARG, "%s: blocks per axis", path);
    R.skip();
    std::vector<char> seen(ndoms, 0);
    for (int q = 0; q < ndoms; ++q) {
      int id = 0, rdiv = 0;
      double c[6];
      if (!R.read(8, t) || !to_int(t[0], id) || !to_int(t[1], rdiv) || id < 0 || id >= ndoms)
        return lfail(H3DMG_ERR_ARG, "%s: domain line %d", path, q + 1);
      for (int e = 0; e < 6; ++e)
        if (!to_double(t[2 + e], c[e]))
          return lfail(H3DMG_ERR_ARG, "%s: domain %d coordinates", path, id);
      Domain &D = L.dom[id];
      D.rdiv = rdiv;
      for (int a = 0; a < 3; ++a) {
        D.lo[a] = c[2 * a];
        D.hi[a] = c[2 * a + 1];
      }
      seen[id] = 1;
    }
    for (int i = 0; i < ndoms; ++i)
      if (!seen[i]) return lfail(H3DMG_ERR_ARG, "%s: domain %d not listed", path, i);
  } else {
    double len[3];
    int nd[3];
    for (int a = 0; a < 3; ++a)
      if (!R.read(1, t) || !to_double(t[0], len[a]))
        return lfail(H3DMG_ERR_ARG, "%s: domain length", path);
    for (int a = 0; a < 3; ++a)
      if (!R.read(1, t) || !to_int(t[0], nd[a]) || nd[a] < 1)
        return lfail(H3DMG_ERR_ARG, "%s: blocks per axis", path);
    if (nd[0] * nd[1] * nd[2] != ndoms)
      return lfail(H3DMG_ERR_ARG, "%s: %d x %d x %d blocks for %d domains", path, nd[0], nd[1],
                   nd[2], ndoms);
    int id = 0;
    for (int k = 0; k < nd[2]; ++k)
      for (int j = 0; j < nd[1]; ++j)
        for (int i = 0; i < nd[0]; ++i, ++id) {
          const int ix[3] = {i, j, k};
          for (int a = 0; a < 3; ++a) {
            const double d = len[a] / (double)nd[a];
            L.dom[id].lo[a] = d * ix[a];
            L.dom[id].hi[a] = d * (ix[a] + 1);
          }
        }
  }
  return H3DMG_OK;
}

constexpr double TOUCH = 1e-8;  // alloc_dom.for's coordinate tolerance

bool close(double a, double b) { return std::fabs(a - b) < TOUCH; }

// the face neighbour of domain d across face f (alloc_dom.for:524-690): the
// domain whose opposite face coincides with it over the same extent along
// the two other axes; -1 for none, -2 when several match (the reference
// stops)
int face_neighbour(const h3dmg_layout &L, int d, int f) {
  const int a = f >> 1;
  const bool next = f & 1;
  const Domain &D = L.dom[d];
  int found = -1, count = 0;
  for (int q = 0; q < (int)L.dom.size(); ++q) {
    if (q == d) continue;
    const Domain &Q = L.dom[q];
    if (!close(next ? D.hi[a] : D.lo[a], next ? Q.lo[a] : Q.hi[a])) continue;
    bool same = true;
    for (int e = 0; e < 3; ++e)
      if (e != a && !(close(D.lo[e], Q.lo[e]) && close(D.hi[e], Q.hi[e]))) same = false;
    if (same) {
      found = q;
      ++count;
    }
  }
  return count > 1 ? -2 : found;
}

// the periodic partner across face f of a domain at the box's boundary on a
// periodic axis (alloc_dom.for:1040-1215): the domain at the opposite end of
// the box over the same extent along the two other axes
int periodic_partner(const h3dmg_layout &L, int d, int f) {
  const int a = f >> 1;
  const bool next = f & 1;
  double st = L.dom[0].lo[a], en = L.dom[0].hi[a];
  for (const Domain &Q : L.dom) {
    st = std::fmin(st, Q.lo[a]);
    en = std::fmax(en, Q.hi[a]);
  }
  const Domain &D = L.dom[d];
  int found = -1, count = 0;
  for (int q = 0; q < (int)L.dom.size(); ++q) {
    if (q == d) continue;
    const Domain &Q = L.dom[q];
    if (!close(next ? st : en, next ? Q.lo[a] : Q.hi[a])) continue;
    bool same = true;
    for (int e = 0; e < 3; ++e)
      if (e != a && !(close(D.lo[e], Q.lo[e]) && close(D.hi[e], Q.hi[e]))) same = false;
    if (same) {
      found = q;
      ++count;
    }
  }
  return count > 1 ? -2 : found;
}

}  // namespace

extern "C" {

const char *h3dmg_layout_error(void) { return g_layout_error.c_str(); }

h3dmg_status h3dmg_layout_read(const char *infodom_path, const char *mdmap_path,
                               h3dmg_layout **out) {
  if (!out || !infodom_path || !mdmap_path) return lfail(H3DMG_ERR_ARG, "null argument");
  *out = nullptr;
  h3dmg_layout *L = new h3dmg_layout();
  h3dmg_status st = read_infodom(infodom_path, *L);
  if (st == H3DMG_OK) st = read_mdmap(mdmap_path, *L);
  if (st != H3DMG_OK) {
    delete L;
    return st;
  }
  *out = L;
  return H3DMG_OK;
}

void h3dmg_layout_destroy(h3dmg_layout *L) { delete L; }

int h3dmg_layout_ndomains(const h3dmg_layout *L) { return L ? (int)L->dom.size() : -1; }
int h3dmg_layout_nranks(const h3dmg_layout *L) { return L ? L->nranks : -1; }

h3dmg_status h3dmg_layout_domain(const h3dmg_layout *L, int d, int *rank, int *rdiv, double *lo,
                                 double *hi) {
  if (!L || d < 0 || d >= (int)L->dom.size()) return lfail(H3DMG_ERR_ARG, "bad domain %d", d);
  const Domain &D = L->dom[d];
  if (rank) *rank = D.rank;
  if (rdiv) *rdiv = D.rdiv;
  for (int a = 0; a < 3; ++a) {
    if (lo) lo[a] = D.lo[a];
    if (hi) hi[a] = D.hi[a];
  }
  return H3DMG_OK;
}

h3dmg_status h3dmg_layout_neighbours(const h3dmg_layout *L, int d, const int *bc_codes,
                                     int *prev_next, int *partner) {
  if (!L || d < 0 || d >= (int)L->dom.size() || !bc_codes)
    return lfail(H3DMG_ERR_ARG, "bad domain %d", d);
  for (int f = 0; f < 6; ++f) {
    const int n = face_neighbour(*L, d, f);
    if (n == -2)
      return lfail(H3DMG_ERR_ARG, "domain %d has more than one neighbour across face %d", d, f);
    if (prev_next) prev_next[f] = n;
  }
  for (int f = 0; f < 6; ++f) {
    int p = -1;
    // per_* is searched where the face is periodic and has no neighbour
    if (bc_codes[f] == 5 && face_neighbour(*L, d, f) < 0) {
      p = periodic_partner(*L, d, f);
      if (p == -2)
        return lfail(H3DMG_ERR_ARG, "domain %d has more than one periodic partner across face %d",
                     d, f);
    }
    if (partner) partner[f] = p;
  }
  return H3DMG_OK;
}

h3dmg_status h3dmg_layout_blocks(const h3dmg_layout *L, int rank, const double *gdx,
                                 const int *bc_codes, int nlevels, h3dmg_block *blocks, int cap,
                                 int *nblocks, int *dom_ids) {
  if (!L || !gdx || !bc_codes || rank < 0 || rank >= L->nranks)
    return lfail(H3DMG_ERR_ARG, "bad layout query (rank %d)", rank);
  const std::vector<int> &mine = L->by_rank[rank];
  if (nblocks) *nblocks = (int)mine.size();
  if (!blocks) return H3DMG_OK;
  if (cap < (int)mine.size())
    return lfail(H3DMG_ERR_ARG, "rank %d has %zu blocks, room for %d", rank, mine.size(), cap);
  std::vector<int> local(L->dom.size(), -1);
  for (size_t q = 0; q < mine.size(); ++q) local[mine[q]] = (int)q;
  for (size_t q = 0; q < mine.size(); ++q) {
    const int d = mine[q];
    const Domain &D = L->dom[d];
    if (D.rdiv != 1)
      return lfail(H3DMG_ERR_UNSUPPORTED, "domain %d: local mesh refinement (rdiv %d)", d,
                   D.rdiv);
    h3dmg_block &B = blocks[q];
    std::memset(&B, 0, sizeof(B));
    int n[3];
    for (int a = 0; a < 3; ++a) {
      // localparameters.for:80-104: cells = extent / spacing, an integer
      const double c = (D.hi[a] - D.lo[a]) / gdx[a];
      if (std::fabs(c - std::nearbyint(c)) > 1e-8 || c < 0.5)
        return lfail(H3DMG_ERR_ARG, "domain %d: extent %g along axis %d is not a multiple of %g",
                     d, D.hi[a] - D.lo[a], a, gdx[a]);
      n[a] = (int)std::nearbyint(c);
    }
    B.ni = n[0];
    B.nj = n[1];
    B.nk = n[2];
    B.dx = gdx[0];
    B.dy = gdx[1];
    B.dz = gdx[2];
    B.nlevels = nlevels;
    int pn[6], pp[6];
    h3dmg_status st = h3dmg_layout_neighbours(L, d, bc_codes, pn, pp);
    if (st) return st;
    for (int f = 0; f < 6; ++f) {
      B.nbr[f] = -1;
      int across = pn[f];
      if (across < 0 && bc_codes[f] == 5) {
        // periodic: onto itself when the block spans the axis, else the
        // partner (exchbc_mgpp, exchangepp.for:2866-2881 / 2884-2905)
        if (pn[f ^ 1] < 0) {
          B.bc[f] = H3DMG_BC_PERIODIC;
          continue;
        }
        across = pp[f];
        if (across < 0)
          return lfail(H3DMG_ERR_ARG, "domain %d: periodic face %d without a partner", d, f);
      }
      if (across < 0) {
        B.bc[f] = H3DMG_BC_WALL;
        continue;
      }
      if (local[across] < 0)
        return lfail(H3DMG_ERR_UNSUPPORTED,
                     "domain %d: neighbour %d is on rank %d (one handle holds one rank's "
                     "blocks; the exchange between ranks is not on the device path)",
                     d, across, L->dom[across].rank);
      B.bc[f] = H3DMG_BC_NEIGHBOUR;
      B.nbr[f] = local[across];
    }
    if (dom_ids) dom_ids[q] = d;
  }
  return H3DMG_OK;
}

}  // extern "C"
