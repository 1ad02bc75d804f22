// fcp_io.hip -- the reference's mesh readers (read_mesh, geometry.f90:90-116)
// on the host side of libfcp: the native polyMesh format
// (read_mesh_native, :118-406) and the OpenFOAM polyMesh format
// (read_mesh_openfoam, :846-1300), parsed into the arrays fcp_mesh_upload
// takes.  Text is parsed with strtod/strtol from one in-memory copy of each
// file.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fcp_internal.h"

namespace {

bool slurp(const std::string &path, std::string &out) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return false;
  fseek(f, 0, SEEK_END);
  const long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  out.resize(sz > 0 ? (size_t)sz : 0);
  const size_t got = sz > 0 ? fread(&out[0], 1, (size_t)sz, f) : 0;
  fclose(f);
  return got == out.size();
}

// Numbers separated by white space (and, in OpenFOAM lists, parentheses).
struct Cursor {
  const char *p, *e;
  explicit Cursor(const std::string &s) : p(s.data()), e(s.data() + s.size()) {}
  void skip(const char *extra = "") {
    while (p < e && (isspace((unsigned char)*p) || (*extra && strchr(extra, *p)))) ++p;
  }
  bool num(long *v, const char *extra = "") {
    skip(extra);
    if (p >= e) return false;
    char *q = nullptr;
    const long r = strtol(p, &q, 10);
    if (q == p) return false;
    p = q;
    *v = r;
    return true;
  }
  bool num(double *v, const char *extra = "") {
    skip(extra);
    if (p >= e) return false;
    char *q = nullptr;
    const double r = strtod(p, &q);
    if (q == p) return false;
    p = q;
    *v = r;
    return true;
  }
  bool expect(char c) {
    skip();
    if (p < e && *p == c) {
      ++p;
      return true;
    }
    return false;
  }
};

// Boundary-type names of polyMesh/boundary (geometry.f90:234-261,
// temperature.f90:168-320); other names carry no treatment.
int bctype_code(const std::string &t) {
  static const char *names[] = {"wall", "symmetry", "inlet", "outlet",
                                "pressure", "empty", "periodic", "wallQFlux"};
  for (int k = 0; k < 8; ++k)
    if (t == names[k]) return k + 1;
  return 0;
}

// The table "name type nFaces startFace [startFaceTwin]" after a header
// line; lines starting with '#' are comments (file_row_count,
// utils.f90:281-360).  A 'periodic' line carries its twin's start face
// (geometry.f90:251-258), collected in patch order.
bool read_boundary(const std::string &path, std::vector<int> &nf, std::vector<int> &sf,
                   std::vector<int> &type, std::vector<int> &twin) {
  std::string txt;
  if (!slurp(path, txt)) return false;
  size_t pos = 0;
  while (pos < txt.size()) {
    size_t eol = txt.find('\n', pos);
    if (eol == std::string::npos) eol = txt.size();
    std::string line = txt.substr(pos, eol - pos);
    pos = eol + 1;
    size_t a = line.find_first_not_of(" \t\r");
    if (a == std::string::npos || line[a] == '#') continue;
    char name[256], tname[256];
    long n = 0, s = 0, t = 0;
    const int got = sscanf(line.c_str() + a, "%255s %255s %ld %ld %ld", name, tname, &n, &s, &t);
    if (got < 4) return false;
    nf.push_back((int)n);
    sf.push_back((int)s);
    type.push_back(bctype_code(tname));
    if (type.back() == FCP_BC_PERIODIC) {
      if (got < 5) return false;
      twin.push_back((int)t);
    }
  }
  return true;
}

// OpenFOAM text without comments and without the FoamFile dictionary.
std::string foam_body(const std::string &txt) {
  std::string out;
  out.reserve(txt.size());
  for (size_t i = 0; i < txt.size();) {
    if (txt.compare(i, 2, "//") == 0) {
      while (i < txt.size() && txt[i] != '\n') ++i;
    } else if (txt.compare(i, 2, "/*") == 0) {
      const size_t j = txt.find("*/", i + 2);
      i = j == std::string::npos ? txt.size() : j + 2;
    } else {
      out += txt[i++];
    }
  }
  const size_t h = out.find("FoamFile");
  if (h != std::string::npos) {
    const size_t c = out.find('}', h);
    if (c != std::string::npos) out.erase(0, c + 1);
  }
  return out;
}

// Count of the first list of an OpenFOAM file and a cursor just inside its '('.
bool foam_list(const std::string &path, std::string &body, long *count, Cursor **cur) {
  std::string txt;
  if (!slurp(path, txt)) return false;
  body = foam_body(txt);
  Cursor *c = new Cursor(body);
  if (!c->num(count) || !c->expect('(')) {
    delete c;
    return false;
  }
  *cur = c;
  return true;
}

}  // namespace

extern "C" {

int fcp_read_polymesh(fcp_ctx *ctx, const char *case_dir, int len, int format, int *sizes) {
  if (!ctx) return FCP_ERR_ARG;
  if (!case_dir || (format != 0 && format != 1))
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_read_polymesh: bad argument");
  std::string dir(case_dir, case_dir + (len > 0 ? len : strlen(case_dir)));
  while (!dir.empty() && dir.back() == ' ') dir.pop_back();
  const std::string d = (dir.empty() ? std::string(".") : dir) + "/polyMesh/";
  std::vector<double> xyz;
  std::vector<int> nnodes, node, owner, neighbour, nf, sf, type, twin;
  long nNodes = 0, nCells = 0, nInner = 0, nBnd = 0, nFaces = 0;
  if (!read_boundary(d + "boundary", nf, sf, type, twin))
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_read_polymesh: cannot read %sboundary", d.c_str());
  auto bad = [&](const char *what) {
    return fcp_fail(ctx, FCP_ERR_ARG, "fcp_read_polymesh: malformed %s%s", d.c_str(), what);
  };
  if (format == 0) {
    std::string txt;
    if (!slurp(d + "size", txt)) return bad("size");
    Cursor c(txt);
    if (!c.num(&nNodes) || !c.num(&nCells) || !c.num(&nInner) || !c.num(&nBnd) || !c.num(&nFaces) ||
        nFaces != nInner + nBnd || nNodes <= 0 || nCells <= 0)
      return bad("size");  // geometry.f90:268-279
    if (!slurp(d + "points", txt)) return bad("points");
    Cursor pc(txt);
    xyz.resize(3 * (size_t)nNodes);
    for (size_t k = 0; k < xyz.size(); ++k)
      if (!pc.num(&xyz[k], ",")) return bad("points");
    if (!slurp(d + "faces", txt)) return bad("faces");
    Cursor fc(txt);
    nnodes.resize(nFaces);
    node.assign((size_t)FCP_NOMAX * nFaces, 0);
    for (long f = 0; f < nFaces; ++f) {
      long nn = 0;
      if (!fc.num(&nn) || nn < 3 || nn > FCP_NOMAX) return bad("faces");
      nnodes[f] = (int)nn;
      for (long k = 0; k < nn; ++k) {
        long v = 0;
        if (!fc.num(&v)) return bad("faces");
        node[(size_t)FCP_NOMAX * f + k] = (int)v;
      }
    }
    if (!slurp(d + "owner", txt)) return bad("owner");
    Cursor oc(txt);
    owner.resize(nFaces);
    for (long f = 0; f < nFaces; ++f) {
      long v = 0;
      if (!oc.num(&v)) return bad("owner");
      owner[f] = (int)v;
    }
    if (!slurp(d + "neighbour", txt)) return bad("neighbour");
    Cursor nc(txt);
    neighbour.resize(nInner);
    for (long f = 0; f < nInner; ++f) {
      long v = 0;
      if (!nc.num(&v)) return bad("neighbour");
      neighbour[f] = (int)v;
    }
  } else {
    std::string body;
    Cursor *c = nullptr;
    long cnt = 0;
    if (!foam_list(d + "points", body, &cnt, &c)) return bad("points");
    nNodes = cnt;
    xyz.resize(3 * (size_t)cnt);
    for (size_t k = 0; k < xyz.size(); ++k)
      if (!c->num(&xyz[k], "()")) {
        delete c;
        return bad("points");
      }
    delete c;
    if (!foam_list(d + "faces", body, &cnt, &c)) return bad("faces");
    nFaces = cnt;
    nnodes.resize(cnt);
    node.assign((size_t)FCP_NOMAX * cnt, 0);
    for (long f = 0; f < cnt; ++f) {
      long nn = 0;
      if (!c->num(&nn, ")") || nn < 3 || nn > FCP_NOMAX || !c->expect('(')) {
        delete c;
        return bad("faces");
      }
      nnodes[f] = (int)nn;
      for (long k = 0; k < nn; ++k) {
        long v = 0;
        if (!c->num(&v)) {
          delete c;
          return bad("faces");
        }
        node[(size_t)FCP_NOMAX * f + k] = (int)v + 1;  // 0-based in OpenFOAM
      }
      if (!c->expect(')')) {
        delete c;
        return bad("faces");
      }
    }
    delete c;
    if (!foam_list(d + "owner", body, &cnt, &c) || cnt != nFaces) return bad("owner");
    owner.resize(cnt);
    long maxc = -1;
    for (long f = 0; f < cnt; ++f) {
      long v = 0;
      if (!c->num(&v)) {
        delete c;
        return bad("owner");
      }
      owner[f] = (int)v + 1;
      maxc = v > maxc ? v : maxc;
    }
    delete c;
    if (!foam_list(d + "neighbour", body, &cnt, &c)) return bad("neighbour");
    nInner = cnt;
    neighbour.resize(cnt);
    for (long f = 0; f < cnt; ++f) {
      long v = 0;
      if (!c->num(&v)) {
        delete c;
        return bad("neighbour");
      }
      neighbour[f] = (int)v + 1;
      maxc = v > maxc ? v : maxc;
    }
    delete c;
    nCells = maxc + 1;
    nBnd = nFaces - nInner;
  }
  const int np = (int)nf.size();
  const int rc = fcp_mesh_upload(ctx, (int)nNodes, (int)nCells, (int)nInner, (int)nBnd, np,
                                 xyz.data(), nnodes.data(), node.data(), owner.data(),
                                 nInner > 0 ? neighbour.data() : nullptr, nf.data(), sf.data(),
                                 type.data());
  if (rc < 0) return rc;
  if (!twin.empty()) FCP_TRY(fcp_mesh_periodic(ctx, (int)twin.size(), twin.data()));
  if (sizes) {
    sizes[0] = (int)nNodes;
    sizes[1] = (int)nCells;
    sizes[2] = (int)nInner;
    sizes[3] = (int)nBnd;
    sizes[4] = np;
  }
  return 0;
}

int fcp_mesh_patches(fcp_ctx *ctx, int *nfaces, int *startFace, int *bctype) {
  if (!ctx) return FCP_ERR_ARG;
  if (!ctx->have_mesh) return fcp_fail(ctx, FCP_ERR_STATE, "fcp_mesh_patches: no mesh");
  for (int ib = 0; ib < ctx->nPatches; ++ib) {
    if (nfaces) nfaces[ib] = ctx->h_nfaces[ib];
    if (startFace) startFace[ib] = ctx->h_start[ib];
    if (bctype) bctype[ib] = ctx->h_bctype[ib];
  }
  return 0;
}

}  // extern "C"
