// hpm_host.cpp -- formats, planar-prior construction, edge map, the HPM facade
// and the pass driver (see hpm_host.hpp).
//
// Float arithmetic follows the reference's expressions: int and float operands
// stay float, double literals promote (C++ rules), pow(x, 2) of a float is x*x;
// built with -ffp-contract=off.
#include "hpm_host.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <thread>
#include <cmath>
#include <mutex>
#include <unordered_map>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <memory>
#include <sstream>

namespace hpmhost {

// ------------------------------------------------------------------ formats

std::vector<Problem> ReadPairFile(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw Error("cannot open " + path);
  int n = 0;
  f >> n;
  std::vector<Problem> out;
  for (int i = 0; i < n && f; ++i) {
    Problem p;
    int m = 0;
    f >> p.ref_image_id >> m;
    for (int j = 0; j < m; ++j) {
      int id;
      float score;
      f >> id >> score;
      if (score <= 0.0f) continue;  // main.cpp:193-195
      p.src_image_ids.push_back(id);
    }
    out.push_back(p);
  }
  if (!f) throw Error("truncated pair file " + path);
  return out;
}

hpm_camera ReadCamera(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw Error("cannot open " + path);
  hpm_camera c;
  std::memset(&c, 0, sizeof(c));
  std::string word;
  f >> word;  // "extrinsic"
  for (int i = 0; i < 3; ++i) f >> c.R[3 * i + 0] >> c.R[3 * i + 1] >> c.R[3 * i + 2] >> c.t[i];
  float row4[4];
  f >> row4[0] >> row4[1] >> row4[2] >> row4[3];
  f >> word;  // "intrinsic"
  for (int i = 0; i < 3; ++i) f >> c.K[3 * i + 0] >> c.K[3 * i + 1] >> c.K[3 * i + 2];
  float interval, depth_num;
  f >> c.depth_min >> interval >> depth_num >> c.depth_max;
  if (!f) throw Error("malformed camera file " + path);
  return c;
}

// Bytes from the read position to the end of the file, or -1.
static int64_t BytesLeft(FILE* fp) {
  const off_t here = ftello(fp);
  if (here < 0 || fseeko(fp, 0, SEEK_END) != 0) return -1;
  const off_t end = ftello(fp);
  if (end < 0 || fseeko(fp, here, SEEK_SET) != 0) return -1;
  return (int64_t)(end - here);
}

// A header's w x h x c elements of `elem` bytes fit in what the file holds: a
// corrupt header is refused before anything is allocated for it.
static bool FitsFile(FILE* fp, int64_t w, int64_t h, int64_t c, int64_t elem) {
  const int64_t left = BytesLeft(fp);
  return left >= 0 && w > 0 && h > 0 && c > 0 && w <= left / elem && h <= left / elem / w &&
         c <= left / elem / w / h;
}

bool ReadDmb(const std::string& path, MapF& out) {
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) return false;
  int32_t hdr[4] = {-1, 0, 0, 0};
  bool ok = std::fread(hdr, sizeof(int32_t), 4, fp) == 4 && hdr[0] == 1 && FitsFile(fp, hdr[2], hdr[1], hdr[3], 4);
  if (ok) {
    out = MapF(hdr[2], hdr[1], hdr[3]);
    ok = std::fread(out.data.data(), sizeof(float), out.data.size(), fp) == out.data.size();
  }
  std::fclose(fp);
  return ok;
}

bool WriteDmb(const std::string& path, const MapF& m) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) return false;
  const int32_t hdr[4] = {1, m.height, m.width, m.channels};
  bool ok = std::fwrite(hdr, sizeof(int32_t), 4, fp) == 4 &&
            std::fwrite(m.data.data(), sizeof(float), m.data.size(), fp) == m.data.size();
  return std::fclose(fp) == 0 && ok;
}

static bool next_token(FILE* fp, std::string& tok) {
  tok.clear();
  int ch;
  while ((ch = std::fgetc(fp)) != EOF) {
    if (ch == '#') {
      while ((ch = std::fgetc(fp)) != EOF && ch != '\n') {
      }
      continue;
    }
    if (!std::isspace(ch)) break;
  }
  if (ch == EOF) return false;
  tok.push_back((char)ch);
  while ((ch = std::fgetc(fp)) != EOF && !std::isspace(ch)) tok.push_back((char)ch);
  return true;  // one whitespace byte after the token has been consumed
}

bool ReadNetpbm(const std::string& path, MapU8& out) {
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) return false;
  std::string magic, sw, sh, sm;
  bool ok = next_token(fp, magic) && (magic == "P5" || magic == "P6") && next_token(fp, sw) && next_token(fp, sh) &&
            next_token(fp, sm) && std::atoi(sm.c_str()) == 255;
  if (ok) {
    const long long lw = std::strtoll(sw.c_str(), nullptr, 10), lh = std::strtoll(sh.c_str(), nullptr, 10);
    const int c = magic == "P6" ? 3 : 1;
    ok = lw < (1 << 30) && lh < (1 << 30) && FitsFile(fp, lw, lh, c, 1);
    const int w = (int)lw, h = (int)lh;
    if (ok) {
      out = MapU8(w, h, c);
      ok = std::fread(out.data.data(), 1, out.data.size(), fp) == out.data.size();
    }
  }
  std::fclose(fp);
  return ok;
}

MapF ToGrayFloat(const MapU8& img) {
  MapF g(img.width, img.height);
  for (size_t i = 0, n = (size_t)img.width * img.height; i < n; ++i) {
    if (img.channels == 1) {
      g.data[i] = (float)img.data[i];
    } else {  // R G B; BT.601 weights in Q14 with rounding, as cv::cvtColor(COLOR_BGR2GRAY) on 8-bit images
      const int y = (4899 * img.data[3 * i] + 9617 * img.data[3 * i + 1] + 1868 * img.data[3 * i + 2] + (1 << 13)) >> 14;
      g.data[i] = (float)y;
    }
  }
  return g;
}

template <class T>
static void ResizeLinearImpl(const Map<T>& src, Map<T>& dst, bool round_u8) {
  const double sx = (double)src.width / dst.width, sy = (double)src.height / dst.height;
  std::vector<int> x0(dst.width), y0(dst.height);
  std::vector<float> ax(dst.width), ay(dst.height);
  auto axis = [](int n, int sn, double s, std::vector<int>& i0, std::vector<float>& a) {
    for (int d = 0; d < n; ++d) {
      double f = (d + 0.5) * s - 0.5;
      int i = (int)std::floor(f);
      float w = (float)(f - i);
      if (i < 0) { i = 0; w = 0.0f; }
      if (i >= sn - 1) { i = sn - 1; w = 0.0f; }
      i0[d] = i;
      a[d] = w;
    }
  };
  axis(dst.width, src.width, sx, x0, ax);
  axis(dst.height, src.height, sy, y0, ay);
  for (int y = 0; y < dst.height; ++y)
    for (int x = 0; x < dst.width; ++x)
      for (int c = 0; c < src.channels; ++c) {
        const int xa = x0[x], xb = std::min(xa + 1, src.width - 1);
        const int ya = y0[y], yb = std::min(ya + 1, src.height - 1);
        const float top = (float)src.at(xa, ya, c) + ax[x] * ((float)src.at(xb, ya, c) - (float)src.at(xa, ya, c));
        const float bot = (float)src.at(xa, yb, c) + ax[x] * ((float)src.at(xb, yb, c) - (float)src.at(xa, yb, c));
        const float v = top + ay[y] * (bot - top);
        dst.at(x, y, c) = round_u8 ? (T)std::min(255.0f, std::max(0.0f, std::floor(v + 0.5f))) : (T)v;
      }
}

MapU8 ResizeLinear(const MapU8& src, int w, int h) {
  MapU8 d(w, h, src.channels);
  ResizeLinearImpl(src, d, true);
  return d;
}

MapF ResizeLinear(const MapF& src, int w, int h) {
  MapF d(w, h, src.channels);
  ResizeLinearImpl(src, d, false);
  return d;
}

namespace {
// Records serialised into blocks of about 8 MB, each written with one fwrite.
class BlockWriter {
 public:
  explicit BlockWriter(FILE* fp) : fp_(fp) { buf_.reserve(kBlock); }
  void put(const void* p, size_t n) {
    if (buf_.size() + n > kBlock) flush();
    const char* c = static_cast<const char*>(p);
    buf_.insert(buf_.end(), c, c + n);
  }
  bool finish() {
    flush();
    return ok_;
  }

 private:
  void flush() {
    if (!buf_.empty()) ok_ = ok_ && std::fwrite(buf_.data(), 1, buf_.size(), fp_) == buf_.size();
    buf_.clear();
  }
  static constexpr size_t kBlock = (size_t)8 << 20;
  FILE* fp_;
  std::vector<char> buf_;
  bool ok_ = true;
};
}  // namespace

bool ExportPointCloud(const std::string& path, const std::vector<FusedPoint>& pts) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) return false;
  std::fprintf(fp,
               "ply\nformat binary_little_endian 1.0\nelement vertex %zu\nproperty float x\nproperty float y\n"
               "property float z\nproperty uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n",
               pts.size());
  BlockWriter w(fp);
  for (const FusedPoint& p : pts) {
    float X[3] = {p.coord[0], p.coord[1], p.coord[2]};
    const bool finite = std::isfinite(X[0]) && std::isfinite(X[1]) && std::isfinite(X[2]);
    if (!finite) X[0] = X[1] = X[2] = 0.0f;  // HPM.cpp:477-481
    const char rgb[3] = {(char)(int)p.color[2], (char)(int)p.color[1], (char)(int)p.color[0]};
    w.put(X, sizeof(X));
    w.put(rgb, 3);
  }
  const bool ok = w.finish();
  return std::fclose(fp) == 0 && ok;
}

bool ExportFusedPointCloud(const std::string& path, const std::vector<FusedPoint>& pts) {
  auto norm2 = [](const FusedPoint& p) {
    return (double)p.normal[0] * p.normal[0] + (double)p.normal[1] * p.normal[1] + (double)p.normal[2] * p.normal[2];
  };
  auto valid = [&](const FusedPoint& p) {
    const double n = norm2(p);
    return std::isfinite(p.coord[0]) && std::isfinite(p.coord[1]) && std::isfinite(p.coord[2]) && std::isfinite(n) &&
           n > 1e-20;
  };
  size_t count = 0;
  for (const FusedPoint& p : pts) count += valid(p);
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) return false;
  std::fprintf(fp,
               "ply\nformat binary_little_endian 1.0\nelement vertex %zu\nproperty float x\nproperty float y\n"
               "property float z\nproperty float nx\nproperty float ny\nproperty float nz\nproperty uchar red\n"
               "property uchar green\nproperty uchar blue\nend_header\n",
               count);
  auto channel = [](float v) -> unsigned char {
    return std::isfinite(v) ? (unsigned char)std::max(0.0f, std::min(255.0f, v)) : 0;
  };
  BlockWriter w(fp);
  for (const FusedPoint& p : pts) {
    if (!valid(p)) continue;
    const double len = std::sqrt(norm2(p));
    const float v[6] = {p.coord[0], p.coord[1], p.coord[2], (float)(p.normal[0] / len), (float)(p.normal[1] / len),
                        (float)(p.normal[2] / len)};
    const unsigned char rgb[3] = {channel(p.color[2]), channel(p.color[1]), channel(p.color[0])};
    w.put(v, sizeof(v));
    w.put(rgb, 3);
  }
  const bool ok = w.finish();
  return std::fclose(fp) == 0 && ok;
}

// ------------------------------------------------------------- planar prior

std::vector<Point2i> GetSupportPoints(const MapF& costs, const MapF& confidences, int width, int height) {
  std::vector<Point2i> pts;
  const int step_size = 5;
  for (int col = 0; col < width; col += step_size) {
    for (int row = 0; row < height; row += step_size) {
      float min_cost_no_texture = 2.0f, min_cost_texture = 2.0f;
      Point2i p_no_texture, p_texture;
      const int c_bound = std::min(width, col + step_size), r_bound = std::min(height, row + step_size);
      for (int c = col; c < c_bound; ++c) {
        for (int r = row; r < r_bound; ++r) {
          const float cost = costs.at(c, r), conf = confidences.at(c, r);
          if (cost < 2.0f) {
            const float f_no_texture = cost - conf;
            const float f_texture = (float)((double)cost + 0.2 - (double)conf);
            if (min_cost_no_texture > f_no_texture) { p_no_texture = {c, r}; min_cost_no_texture = f_no_texture; }
            if (min_cost_texture > f_texture) { p_texture = {c, r}; min_cost_texture = f_texture; }
          }
        }
      }
      if ((double)min_cost_texture < 0.1) pts.push_back(p_texture);
      if ((double)min_cost_no_texture < 0.1) pts.push_back(p_no_texture);
    }
  }
  return pts;
}

namespace {

// Incremental Delaunay triangulation (Bowyer-Watson) with exact predicates on
// integer coordinates: orientation in int64, in-circle in __int128.
struct Delaunay {
  struct V { int64_t x, y; };
  struct T { int v[3]; int nb[3]; bool alive; };  // nb[i] is across the edge opposite v[i]
  std::vector<V> vs;
  std::vector<T, MapAllocator<T>> ts;  // about 5 per point; dead ones stay (their order is the output order)
  int last = 0;
  // per-insert scratch kept across inserts (reset by stamp or by the entries
  // touched), so that an insert costs the size of its cavity, not of the mesh
  std::vector<uint32_t> mark;  // mark[t] == stamp: t is in the current cavity
  uint32_t stamp = 0;
  std::vector<int> vstart, vend;  // per vertex: new triangle starting / ending there
  std::vector<int> cavity, stack, created;
  struct E { int a, b, out; };
  std::vector<E> boundary;

  static int64_t orient(const V& a, const V& b, const V& c) {
    return (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x);
  }
  // > 0 when d lies strictly inside the circumcircle of the CCW triangle abc.
  // Exact: with every coordinate difference below 2^14 each term of the
  // determinant is below 2^59 and 64-bit integers suffice; tests that reach
  // the far bounding vertices take 128 bits.
  static int incircle(const V& a, const V& b, const V& c, const V& d) {
    const int64_t adx = a.x - d.x, ady = a.y - d.y, bdx = b.x - d.x, bdy = b.y - d.y, cdx = c.x - d.x,
                  cdy = c.y - d.y;
    const int64_t lim = (int64_t)1 << 14;
    if (std::max({std::abs(adx), std::abs(ady), std::abs(bdx), std::abs(bdy), std::abs(cdx), std::abs(cdy)}) < lim) {
      const int64_t ad = adx * adx + ady * ady, bd = bdx * bdx + bdy * bdy, cd = cdx * cdx + cdy * cdy;
      const int64_t det = adx * (bdy * cd - bd * cdy) - ady * (bdx * cd - bd * cdx) + ad * (bdx * cdy - bdy * cdx);
      return det > 0 ? 1 : (det < 0 ? -1 : 0);
    }
    return incircle128(adx, ady, bdx, bdy, cdx, cdy);
  }
  static int incircle128(__int128 adx, __int128 ady, __int128 bdx, __int128 bdy, __int128 cdx, __int128 cdy) {
    const __int128 ad = adx * adx + ady * ady, bd = bdx * bdx + bdy * bdy, cd = cdx * cdx + cdy * cdy;
    const __int128 det = adx * (bdy * cd - bd * cdy) - ady * (bdx * cd - bd * cdx) + ad * (bdx * cdy - bdy * cdx);
    return det > 0 ? 1 : (det < 0 ? -1 : 0);
  }

  // Coordinates are stored doubled so that the bounding triangle is integral.
  // The bounding triangle is the one cv::Subdiv2D starts from (centre of the
  // rectangle +- 3 x its larger side), so that, as in the reference, thin
  // triangles along the convex hull whose circumcircle reaches a bounding
  // vertex are not produced.
  Delaunay(int64_t w, int64_t h) {
    const int64_t big = 3 * std::max(w, h), cx = w, cy = h;  // doubled centre
    vs = {{cx + 2 * big, cy}, {cx, cy + 2 * big}, {cx - 2 * big, cy - 2 * big}};
    ts.push_back(T{{0, 1, 2}, {-1, -1, -1}, true});
    vstart.assign(3, -1);
    vend.assign(3, -1);
  }

  int locate(const V& p) {
    int t = last;
    if (!ts[t].alive) t = (int)ts.size() - 1;
    for (size_t steps = 0; steps < 4 * ts.size() + 16; ++steps) {
      const T& tr = ts[t];
      int next = -1;
      for (int k = 0; k < 3 && next < 0; ++k) {
        const int i = (k + (int)(steps % 3)) % 3;  // rotate the test order so a walk cannot cycle
        if (orient(vs[tr.v[(i + 1) % 3]], vs[tr.v[(i + 2) % 3]], p) < 0) next = tr.nb[i];
      }
      if (next < 0) return t;
      t = next;
    }
    for (size_t i = 0; i < ts.size(); ++i) {  // fallback: scan
      if (!ts[i].alive) continue;
      const T& tr = ts[i];
      if (orient(vs[tr.v[0]], vs[tr.v[1]], p) >= 0 && orient(vs[tr.v[1]], vs[tr.v[2]], p) >= 0 &&
          orient(vs[tr.v[2]], vs[tr.v[0]], p) >= 0)
        return (int)i;
    }
    return -1;
  }

  void insert(int64_t x, int64_t y) {
    const V p{x, y};
    const int t0 = locate(p);
    if (t0 < 0) return;
    for (int k = 0; k < 3; ++k)
      if (vs[ts[t0].v[k]].x == x && vs[ts[t0].v[k]].y == y) return;  // duplicate point
    const int pi = (int)vs.size();
    vs.push_back(p);
    vstart.push_back(-1);
    vend.push_back(-1);
    if (mark.size() < ts.size()) mark.resize(ts.size() + ts.size() / 2 + 16, 0);
    if (++stamp == 0) {  // wrapped: clear the marks once
      std::fill(mark.begin(), mark.end(), 0);
      stamp = 1;
    }
    // cavity: triangles whose circumcircle strictly contains p, grown from t0
    cavity.assign(1, t0);
    stack.assign(1, t0);
    mark[t0] = stamp;
    while (!stack.empty()) {
      const int t = stack.back();
      stack.pop_back();
      for (int k = 0; k < 3; ++k) {
        const int n = ts[t].nb[k];
        if (n < 0 || mark[n] == stamp) continue;
        const T& tn = ts[n];
        if (incircle(vs[tn.v[0]], vs[tn.v[1]], vs[tn.v[2]], p) > 0) {
          mark[n] = stamp;
          cavity.push_back(n);
          stack.push_back(n);
        }
      }
    }
    // boundary edges (a, b) in CCW order of their cavity triangle, with the
    // outside neighbour; each becomes the triangle (a, b, p)
    boundary.clear();
    for (int t : cavity)
      for (int k = 0; k < 3; ++k) {
        const int n = ts[t].nb[k];
        if (n < 0 || mark[n] != stamp) boundary.push_back({ts[t].v[(k + 1) % 3], ts[t].v[(k + 2) % 3], n});
      }
    for (int t : cavity) ts[t].alive = false;
    created.clear();
    for (const E& e : boundary) {
      const int id = (int)ts.size();
      ts.push_back(T{{e.a, e.b, pi}, {-1, -1, e.out}, true});
      created.push_back(id);
      vstart[e.a] = id;
      vend[e.b] = id;
      if (e.out >= 0)
        for (int k = 0; k < 3; ++k) {
          T& o = ts[e.out];
          // the outside triangle's edge (b, a) is opposite its vertex k
          if (o.v[(k + 1) % 3] == e.b && o.v[(k + 2) % 3] == e.a) o.nb[k] = id;
        }
    }
    for (int id : created) {
      T& t = ts[id];
      t.nb[0] = vstart[t.v[1]];  // edge (b, p) is shared with the triangle that starts at b
      t.nb[1] = vend[t.v[0]];    // edge (p, a) with the triangle that ends at a
    }
    for (const E& e : boundary) vstart[e.a] = vend[e.b] = -1;
    last = created.empty() ? 0 : created.back();
  }
};

}  // namespace

std::vector<Triangle> DelaunayTriangulation(int width, int height, const std::vector<Point2i>& points) {
  std::vector<Triangle> out;
  if (points.empty()) return out;
  Delaunay d(width, height);
  // room for the triangles an incremental build allocates (about 5 per
  // point), so that the arrays are not copied while they grow
  d.ts.reserve(5 * points.size() + 64);
  d.mark.resize(5 * points.size() + 64, 0);
  d.vs.reserve(points.size() + 3);
  d.vstart.reserve(points.size() + 3);
  d.vend.reserve(points.size() + 3);
  for (const Point2i& p : points)
    if (p.x >= 0 && p.x < width && p.y >= 0 && p.y < height) d.insert(2 * (int64_t)p.x, 2 * (int64_t)p.y);
  for (const auto& t : d.ts) {
    if (!t.alive || t.v[0] < 3 || t.v[1] < 3 || t.v[2] < 3) continue;
    const auto& a = d.vs[t.v[0]];
    const auto& b = d.vs[t.v[1]];
    const auto& c = d.vs[t.v[2]];
    out.push_back(Triangle{{(int)(a.x / 2), (int)(a.y / 2)}, {(int)(b.x / 2), (int)(b.y / 2)},
                           {(int)(c.x / 2), (int)(c.y / 2)}});
  }
  return out;
}

void RasterTriangles(const std::vector<Triangle>& tris, MapF& mask) {
  uint32_t idx = 0;
  for (const Triangle& t : tris) {
    auto in = [&](const Point2i& p) { return p.x >= 0 && p.x < mask.width && p.y >= 0 && p.y < mask.height; };
    if (!(in(t.pt1) && in(t.pt2) && in(t.pt3))) continue;
    auto len = [](const Point2i& a, const Point2i& b) {
      const double dx = a.x - b.x, dy = a.y - b.y;  // pow(int, 2) is a double
      return (float)std::sqrt(dx * dx + dy * dy);
    };
    const float max_edge = std::max(len(t.pt1, t.pt2), std::max(len(t.pt1, t.pt3), len(t.pt2, t.pt3)));
    const float step = (float)(1.0 / max_edge);
    for (float p = 0; p < 1.0; p += step) {
      for (float q = 0; (double)q < 1.0 - p; q += step) {
        const int x = (int)((double)(p * t.pt1.x + q * t.pt2.x) + (1.0 - p - q) * t.pt3.x);
        const int y = (int)((double)(p * t.pt1.y + q * t.pt2.y) + (1.0 - p - q) * t.pt3.y);
        if (x >= 0 && x < mask.width && y >= 0 && y < mask.height) mask.at(x, y) = (float)(idx + 1.0);
      }
    }
    ++idx;
  }
}

void PlaneThroughPoints(const float X1[3], const float X2[3], const float X3[3], float plane[4]) {
  // Null vector of [X 1] (3 x 4): the generalised cross product of its rows,
  // component k = (-1)^k det(rows without column k), in double.
  const double A[3][4] = {{X1[0], X1[1], X1[2], 1.0}, {X2[0], X2[1], X2[2], 1.0}, {X3[0], X3[1], X3[2], 1.0}};
  double n[4];
  for (int k = 0; k < 4; ++k) {
    int c[3], m = 0;
    for (int j = 0; j < 4; ++j)
      if (j != k) c[m++] = j;
    const double det = A[0][c[0]] * (A[1][c[1]] * A[2][c[2]] - A[1][c[2]] * A[2][c[1]]) -
                       A[0][c[1]] * (A[1][c[0]] * A[2][c[2]] - A[1][c[2]] * A[2][c[0]]) +
                       A[0][c[2]] * (A[1][c[0]] * A[2][c[1]] - A[1][c[1]] * A[2][c[0]]);
    n[k] = (k % 2 ? -det : det);
  }
  const double s = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2] + n[3] * n[3]);
  float n4[4];
  for (int k = 0; k < 4; ++k) n4[k] = (float)(n[k] / s);  // unit null vector, as SVD::solveZ returns
  float norm2 = (float)std::sqrt((double)n4[0] * n4[0] + (double)n4[1] * n4[1] + (double)n4[2] * n4[2]);
  if (n4[3] < 0) norm2 *= -1;
  for (int k = 0; k < 4; ++k) plane[k] = n4[k] / norm2;
}

// ----------------------------------------------------------------- edge map

MapU8 CannyEdges(const MapU8& gray, float low, float high) {
  const int W = gray.width, H = gray.height;
  const int C = gray.channels;
  if (C != 1 && C != 3) throw Error("Canny: 1 or 3 channels expected");
  auto px = [&](int x, int y, int c) {
    x = std::min(std::max(x, 0), W - 1);
    y = std::min(std::max(y, 0), H - 1);
    return (int)gray.data[((size_t)y * W + x) * C + c];
  };
  std::vector<int> gx((size_t)W * H), gy((size_t)W * H), mag((size_t)W * H, -1);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const size_t i = (size_t)y * W + x;
      // a colour image keeps, per pixel, the channel of largest |dx| + |dy|;
      // channels are visited B, G, R as in an OpenCV BGR image, first wins ties
      for (int k = 0; k < C; ++k) {
        const int c = C - 1 - k;
        const int dx = (px(x + 1, y - 1, c) + 2 * px(x + 1, y, c) + px(x + 1, y + 1, c)) -
                       (px(x - 1, y - 1, c) + 2 * px(x - 1, y, c) + px(x - 1, y + 1, c));
        const int dy = (px(x - 1, y + 1, c) + 2 * px(x, y + 1, c) + px(x + 1, y + 1, c)) -
                       (px(x - 1, y - 1, c) + 2 * px(x, y - 1, c) + px(x + 1, y - 1, c));
        const int v = std::abs(dx) + std::abs(dy);
        if (v > mag[i]) {
          gx[i] = dx;
          gy[i] = dy;
          mag[i] = v;
        }
      }
    }
  auto m = [&](int x, int y) { return (x < 0 || y < 0 || x >= W || y >= H) ? 0 : mag[(size_t)y * W + x]; };
  // 0 none, 1 weak, 2 strong
  std::vector<uint8_t> cls((size_t)W * H, 0);
  const double t22 = 0.41421356237309503, t67 = 2.414213562373095;  // tan 22.5, tan 67.5 degrees
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const size_t i = (size_t)y * W + x;
      const int v = mag[i];
      if (v <= low) continue;
      const double ax = std::abs(gx[i]), ay = std::abs(gy[i]);
      bool peak;
      if (ay <= ax * t22) peak = v > m(x - 1, y) && v >= m(x + 1, y);
      else if (ay >= ax * t67) peak = v > m(x, y - 1) && v >= m(x, y + 1);
      else if ((gx[i] ^ gy[i]) < 0) peak = v > m(x + 1, y - 1) && v >= m(x - 1, y + 1);
      else peak = v > m(x - 1, y - 1) && v >= m(x + 1, y + 1);
      if (peak) cls[i] = v > high ? 2 : 1;
    }
  MapU8 out(W, H);
  std::vector<size_t> stack;
  for (size_t i = 0; i < cls.size(); ++i)
    if (cls[i] == 2) { out.data[i] = 255; stack.push_back(i); }
  while (!stack.empty()) {
    const size_t i = stack.back();
    stack.pop_back();
    const int x = (int)(i % W), y = (int)(i / W);
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int xx = x + dx, yy = y + dy;
        if (xx < 0 || yy < 0 || xx >= W || yy >= H) continue;
        const size_t j = (size_t)yy * W + xx;
        if (cls[j] == 1 && !out.data[j]) { out.data[j] = 255; stack.push_back(j); }
      }
  }
  return out;
}

// --------------------------------------------------------------- HPM facade

std::string ResultFolder(const std::string& dense_folder, int ref_image_id) {
  std::ostringstream s;
  s << dense_folder << "/HPM_MVS_plusplus/2333_" << std::setw(8) << std::setfill('0') << ref_image_id;
  return s.str();
}

static std::string ImagePath(const std::string& dense_folder, int id, const char* ext) {
  std::ostringstream s;
  s << dense_folder << "/images/" << std::setw(8) << std::setfill('0') << id << ext;
  return s.str();
}

// The size from the Netpbm header only.
static void ImageSize(const std::string& dense_folder, int id, int& w, int& h) {
  for (const char* ext : {".pgm", ".ppm"}) {
    FILE* fp = std::fopen(ImagePath(dense_folder, id, ext).c_str(), "rb");
    if (!fp) continue;
    std::string magic, sw, sh;
    const bool ok = next_token(fp, magic) && (magic == "P5" || magic == "P6") && next_token(fp, sw) && next_token(fp, sh);
    std::fclose(fp);
    if (ok && (w = std::atoi(sw.c_str())) > 0 && (h = std::atoi(sh.c_str())) > 0) return;
  }
  if (ReadJpegSize(ImagePath(dense_folder, id, ".jpg"), w, h)) return;
  throw Error("cannot read the size of image " + ImagePath(dense_folder, id, ".pgm|.ppm|.jpg"));
}

static MapU8 LoadImage(const std::string& dense_folder, int id) {
  MapU8 img;
  if (ReadNetpbm(ImagePath(dense_folder, id, ".pgm"), img) || ReadNetpbm(ImagePath(dense_folder, id, ".ppm"), img))
    return img;
  std::string err;
  const std::string jpg = ImagePath(dense_folder, id, ".jpg");
  if (ReadJpeg(jpg, img, &err)) return img;
  throw Error("cannot read image " + ImagePath(dense_folder, id, ".pgm|.ppm|.jpg") + ": " + err);
}

void SceneImages::ensure(int id) {
  if (id < 0) throw Error("negative image id");
  if ((size_t)id >= orig_.size()) {
    orig_.resize(id + 1);
    cur_.resize(id + 1);
    gray_.resize(id + 1);
    edges_.resize(id + 1);
    have_orig_.resize(id + 1, false);
    have_cur_.resize(id + 1, false);
    have_gray_.resize(id + 1, false);
    have_edges_.resize(id + 1, false);
  }
}

SceneImages::~SceneImages() { drop_pool(); }

void SceneImages::drop_pool() {
  if (pool_) hpm_pool_destroy(pool_);
  pool_ = nullptr;
  pool_index_.clear();
}

const MapF& SceneImages::CurrentGray(int id) {
  ensure(id);
  if (!have_gray_[id]) {
    gray_[id] = ToGrayFloat(Current(id));
    have_gray_[id] = true;
  }
  return gray_[id];
}

// cv::Canny(image, 50, 150) of ProcessProblem (main.cpp:280-289), computed on
// device 0 (the reference's device, main.cpp:250); the same map as CannyEdges.
static MapU8 DeviceEdges(const MapU8& img) {
  MapU8 e(img.width, img.height);
  if (hpm_canny(0, img.data.data(), img.width, img.height, img.channels, 50.0f, 150.0f, e.data.data()) != HPM_OK)
    throw Error("hpm_canny failed");
  return e;
}

const MapU8& SceneImages::CurrentEdges(int id) {
  ensure(id);
  if (!have_edges_[id]) {
    edges_[id] = DeviceEdges(Current(id));
    have_edges_[id] = true;
  }
  return edges_[id];
}

hpm_camera SceneImages::CurrentCamera(int id) {
  std::ostringstream cp;
  cp << folder_ << "/cams/" << std::setw(8) << std::setfill('0') << id << "_cam.txt";
  hpm_camera c = ReadCamera(cp.str());
  const MapU8& img = Current(id);
  // K scaled by the prepared / original size, HPM.cpp:668-674
  const MapU8& orig = Original(id);
  const float scale_x = img.width / static_cast<float>(orig.width);
  const float scale_y = img.height / static_cast<float>(orig.height);
  c.K[0] *= scale_x; c.K[2] *= scale_x;
  c.K[4] *= scale_y; c.K[5] *= scale_y;
  c.width = img.width;
  c.height = img.height;
  return c;
}

void SceneImages::BuildPool(const std::vector<int>& ids, int device) {
  drop_pool();
  std::vector<hpm_camera> cams;
  std::vector<const float*> ptrs;
  for (int id : ids) {
    cams.push_back(CurrentCamera(id));
    ptrs.push_back(CurrentGray(id).data.data());
    CurrentEdges(id);  // derived here, so that the passes' threads only read the cache
  }
  if (hpm_pool_create(&pool_, device, (int)ids.size(), cams.data(), ptrs.data()) != HPM_OK) {
    pool_ = nullptr;
    throw Error("hpm_pool_create failed");
  }
  for (size_t k = 0; k < ids.size(); ++k) {
    if ((size_t)ids[k] >= pool_index_.size()) pool_index_.resize(ids[k] + 1, -1);
    pool_index_[ids[k]] = (int)k;
  }
}

int SceneImages::PoolIndex(int id) const {
  return (pool_ && id >= 0 && (size_t)id < pool_index_.size()) ? pool_index_[id] : -1;
}

const MapU8& SceneImages::Original(int id) {
  ensure(id);
  if (!have_orig_[id]) {
    orig_[id] = LoadImage(folder_, id);
    have_orig_[id] = true;
  }
  return orig_[id];
}

void SceneImages::PrepareLevel(int id, int target_max_side) {  // main.cpp:123-145
  drop_pool();
  const MapU8& img = Original(id);
  have_gray_[id] = false;
  have_edges_[id] = false;
  if (std::max(img.width, img.height) > target_max_side) {
    const float factor = std::min((float)target_max_side / img.width, (float)target_max_side / img.height);
    const int nw = std::max(1, (int)std::round(img.width * factor));
    const int nh = std::max(1, (int)std::round(img.height * factor));
    cur_[id] = ResizeLinear(img, nw, nh);
  } else {
    cur_[id] = img;
  }
  have_cur_[id] = true;
}

const MapU8& SceneImages::Current(int id) {
  ensure(id);
  return have_cur_[id] ? cur_[id] : Original(id);
}

// ------------------------------------------------------------- map store
//
// During a pipeline run every .dmb the passes write is also kept in memory,
// and later passes read it from there: a pass reads many maps the passes
// before it wrote (a geometric view reads the depth maps of all its
// sources).  The files are still written, by a background thread, in the
// order the maps were produced; a map replaced before its file was written
// is written once, in its last version.  The run waits for the writer before
// it returns, so the files are those the passes would have written.  Outside
// a run (single passes through the C ABI) maps are read from and written to
// the files directly.
namespace {
class MapStore {
 public:
  MapStore() : writer_([this] { run(); }) {}
  ~MapStore() {
    {
      std::lock_guard<std::mutex> l(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    writer_.join();
  }
  MapRef get(const std::string& path) {
    std::lock_guard<std::mutex> l(mu_);
    auto it = maps_.find(path);
    return it == maps_.end() ? nullptr : it->second;
  }
  void put(const std::string& path, MapRef m) {
    {
      std::lock_guard<std::mutex> l(mu_);
      maps_[path] = m;
      auto it = pending_.find(path);
      if (it != pending_.end()) {
        it->second = std::move(m);  // still queued: write the last version only
      } else {
        pending_.emplace(path, std::move(m));
        order_.push_back(path);
      }
    }
    cv_.notify_all();
  }
  // A map read from its file: kept for later reads, not written back.
  void put_read(const std::string& path, MapRef m) {
    std::lock_guard<std::mutex> l(mu_);
    maps_.emplace(path, std::move(m));
  }
  // Waits until every queued map is on disk; throws on a failed write.
  void flush() {
    std::unique_lock<std::mutex> l(mu_);
    cv_.wait(l, [this] { return order_.empty() && !busy_; });
    if (!error_.empty()) {
      const std::string e = error_;
      error_.clear();
      throw Error(e);
    }
  }

 private:
  void run() {
    std::unique_lock<std::mutex> l(mu_);
    for (;;) {
      cv_.wait(l, [this] { return stop_ || !order_.empty(); });
      if (order_.empty()) return;  // stop requested and nothing left
      const std::string path = order_.front();
      order_.pop_front();
      MapRef m = std::move(pending_[path]);
      pending_.erase(path);
      busy_ = true;
      l.unlock();
      const bool ok = WriteDmb(path, *m);
      m.reset();
      l.lock();
      busy_ = false;
      if (!ok && error_.empty()) error_ = "cannot write " + path;
      cv_.notify_all();
    }
  }
  std::mutex mu_;
  std::condition_variable cv_;
  std::unordered_map<std::string, MapRef> maps_, pending_;
  std::deque<std::string> order_;
  bool busy_ = false, stop_ = false;
  std::string error_;
  std::thread writer_;
};

std::mutex g_store_mu;
std::shared_ptr<MapStore> g_store;

std::shared_ptr<MapStore> active_store() {
  std::lock_guard<std::mutex> l(g_store_mu);
  return g_store;
}
}  // namespace

MapStoreScope::MapStoreScope() {
  std::lock_guard<std::mutex> l(g_store_mu);
  if (!g_store) {
    g_store = std::make_shared<MapStore>();
    owner_ = true;
  }
}

void MapStoreScope::Finish() {
  if (!owner_) return;
  std::shared_ptr<MapStore> s = active_store();
  if (s) s->flush();
}

MapStoreScope::~MapStoreScope() {
  if (!owner_) return;
  std::shared_ptr<MapStore> s;
  {
    std::lock_guard<std::mutex> l(g_store_mu);
    s = std::move(g_store);
    g_store.reset();
  }
  // the destructor writes what is still queued before it returns
}

MapRef LoadMap(const std::string& path) {
  std::shared_ptr<MapStore> s = active_store();
  if (s) {
    if (MapRef m = s->get(path)) return m;
  }
  auto m = std::make_shared<MapF>();
  if (!ReadDmb(path, *m)) throw Error("cannot read " + path);
  if (s) s->put_read(path, m);
  return m;
}

// A loaded map must have the shape its consumer reads: a .dmb of another size
// or channel count would otherwise be read past its end.
static const MapF& Shaped(const MapRef& m, int w, int h, int c, const std::string& path) {
  if (m->width != w || m->height != h || m->channels != c) {
    std::ostringstream os;
    os << path << ": a " << m->width << "x" << m->height << "x" << m->channels << " map where " << w << "x" << h
       << "x" << c << " is needed";
    throw Error(os.str());
  }
  return *m;
}

void SaveMap(const std::string& path, MapF&& m) {
  std::shared_ptr<MapStore> s = active_store();
  if (!s) {
    if (!WriteDmb(path, m)) throw Error("cannot write " + path);
    return;
  }
  s->put(path, std::make_shared<const MapF>(std::move(m)));
}



HPM::HPM() {
  hpm_default_params(&params_);
  if (hpm_create(&ctx_, 0) != HPM_OK) throw Error("hpm_create failed (no HIP device?)");
}

HPM::~HPM() {
  hpm_batch_destroy(fused_);
  hpm_destroy(ctx_);
}

void HPM::check(int rc, const char* what) const {
  if (rc != HPM_OK) throw Error(std::string(what) + ": " + hpm_last_error(ctx_));
}

void HPM::SetGeomConsistencyParams(bool multi_geometry) {  // HPM.cpp:565-569
  params_.geom_consistency = 1;
  multi_geometry_ = multi_geometry;
}
void HPM::SetPlanarPriorParams() { params_.prior_consistency = 1; }
void HPM::SetHierarchyParams() { params_.hierarchy = 1; }
void HPM::SetMandConsistencyParams(bool flag) { params_.mand_consistency = flag ? 1 : 0; }
void HPM::SetPrecision(int precision) { params_.precision = precision; }

void HPM::InuputInitialization(const std::string& dense_folder, const std::vector<Problem>& problems, int idx,
                               SceneImages* images) {
  images_.clear();
  owned_images_.clear();
  cams_.clear();
  depths_.clear();
  pool_ = nullptr;
  pool_ids_.clear();
  const Problem& problem = problems.at(idx);
  std::vector<int> ids{problem.ref_image_id};
  ids.insert(ids.end(), problem.src_image_ids.begin(), problem.src_image_ids.end());
  if (images) {
    for (int id : ids) {
      cams_.push_back(images->CurrentCamera(id));
      images_.push_back(&images->CurrentGray(id));
      pool_ids_.push_back(images->PoolIndex(id));
    }
    const bool all = std::all_of(pool_ids_.begin(), pool_ids_.end(), [](int k) { return k >= 0; });
    pool_ = all ? images->Pool() : nullptr;
  } else {
    owned_images_.reserve(ids.size());
    for (int id : ids) {
      std::ostringstream cp;
      cp << dense_folder << "/cams/" << std::setw(8) << std::setfill('0') << id << "_cam.txt";
      hpm_camera c = ReadCamera(cp.str());
      const MapU8 img = LoadImage(dense_folder, id);
      c.width = img.width;
      c.height = img.height;
      cams_.push_back(c);
      owned_images_.push_back(ToGrayFloat(img));
    }
    for (const MapF& m : owned_images_) images_.push_back(&m);
  }
  if (params_.geom_consistency) {  // HPM.cpp:693-732
    const char* suffix = multi_geometry_ ? "/depths_geom.dmb" : "/depths.dmb";
    for (size_t k = 0; k < ids.size(); ++k) {
      const std::string path = ResultFolder(dense_folder, ids[k]) + suffix;
      depths_.push_back(LoadMap(path));
      Shaped(depths_.back(), cams_[k].width, cams_[k].height, 1, path);
    }
  }
}

void HPM::CudaSpaceInitialization(const std::string& dense_folder, const Problem& problem) {
  check(hpm_set_params(ctx_, &params_), "hpm_set_params");
  if (pool_) {  // the level's images are on the device already
    check(hpm_set_views_pooled(ctx_, pool_, (int)pool_ids_.size(), pool_ids_.data()), "hpm_set_views_pooled");
  } else {
    std::vector<const float*> ptrs;
    for (const MapF* m : images_) ptrs.push_back(m->data.data());
    check(hpm_set_views(ctx_, (int)cams_.size(), cams_.data(), ptrs.data()), "hpm_set_views");
  }
  if (params_.geom_consistency) {
    std::vector<const float*> dp;
    for (const MapRef& m : depths_) dp.push_back(m->data.data());
    check(hpm_set_src_depths(ctx_, dp.data()), "hpm_set_src_depths");
    // reload of the reference's own maps, HPM.cpp:813-846
    const std::string rf = ResultFolder(dense_folder, problem.ref_image_id);
    CudaHypothesesReload(*depths_[0], *LoadMap(rf + "/costs.dmb"), *LoadMap(rf + "/normals.dmb"));
  }
  if (params_.hierarchy) {  // HPM.cpp:849-905
    const std::string rf = ResultFolder(dense_folder, problem.ref_image_id);
    const MapRef fine_ref = LoadMap(rf + "/depths.dmb");  // the JBU-upsampled depth
    const MapRef cn_ref = LoadMap(rf + "/normals.dmb");   // the coarser level's maps
    const MapRef cc_ref = LoadMap(rf + "/costs.dmb");
    const MapF& fine_depth =
        Shaped(fine_ref, GetReferenceImageWidth(), GetReferenceImageHeight(), 1, rf + "/depths.dmb");
    const int cw = cn_ref->width, ch = cn_ref->height;
    const MapF& cn = Shaped(cn_ref, cw, ch, 3, rf + "/normals.dmb");
    const MapF& cc = Shaped(cc_ref, cw, ch, 1, rf + "/costs.dmb");
    // the reference compares the coarse width with the fine rows (HPM.cpp:869)
    const bool upsample = cw != GetReferenceImageHeight() || ch != GetReferenceImageWidth();
    std::vector<float> coarse((size_t)4 * cw * ch);
    for (size_t i = 0; i < (size_t)cw * ch; ++i) {
      coarse[4 * i + 0] = cn.data[3 * i + 0];
      coarse[4 * i + 1] = cn.data[3 * i + 1];
      coarse[4 * i + 2] = cn.data[3 * i + 2];
      coarse[4 * i + 3] = upsample ? cc.data[i] : (i < fine_depth.data.size() ? fine_depth.data[i] : 0.0f);
    }
    check(hpm_set_coarse(ctx_, coarse.data(), cw, ch, fine_depth.data.data()), "hpm_set_coarse");
  }
}

void HPM::CudaCannyInitialization(const MapU8& canny) {
  if (canny.width != GetReferenceImageWidth() || canny.height != GetReferenceImageHeight() || canny.channels != 1)
    throw Error("edge map of another size than the reference image");
  check(hpm_set_canny_u8(ctx_, canny.data.data()), "hpm_set_canny_u8");
}

void HPM::CudaConfidenceInitialization(const std::string& dense_folder, const std::vector<Problem>& problems, int idx) {
  const std::string path = ResultFolder(dense_folder, problems.at(idx).ref_image_id) + "/confidence.dmb";
  const MapRef conf = LoadMap(path);
  Shaped(conf, GetReferenceImageWidth(), GetReferenceImageHeight(), 1, path);
  check(hpm_set_confidence(ctx_, conf->data.data()), "hpm_set_confidence");
}

void HPM::CudaHypothesesReload(const MapF& depths, const MapF& costs, const MapF& normals) {
  const int w = GetReferenceImageWidth(), h = GetReferenceImageHeight();
  for (const MapF* m : {&depths, &costs, &normals})
    if (m->width != w || m->height != h) throw Error("reload: maps of another size than the reference image");
  if (normals.channels != 3 || depths.channels != 1 || costs.channels != 1) throw Error("reload: bad map channels");
  check(hpm_reload_hypotheses_maps(ctx_, depths.data.data(), normals.data.data(), costs.data.data()),
        "hpm_reload_hypotheses_maps");
}

void HPM::BuildPlanarPrior(int distance) {
  const int width = GetReferenceImageWidth(), height = GetReferenceImageHeight();
  int mw = width, mh = height;
  if (distance > 0) {  // the grid of PriorFromCoarser
    const float factor = (float)(1.0 / (distance * 2));
    mw = (int)std::round(width * factor);
    mh = (int)std::round(height * factor);
  }
  const int cap = 2 * ((width + 4) / 5) * ((height + 4) / 5);
  std::vector<int32_t> xy((size_t)2 * cap);
  int count = 0;
  check(hpm_prior_support_points(ctx_, distance, xy.data(), cap, &count), "hpm_prior_support_points");
  if (count > cap) throw Error("hpm_prior_support_points: more points than blocks");
  std::vector<Point2i> support((size_t)count);
  for (int i = 0; i < count; ++i) support[i] = {xy[2 * i], xy[2 * i + 1]};
  const std::vector<Triangle> tris = hpmhost::DelaunayTriangulation(mw, mh, support);
  std::vector<int32_t> tv(6 * tris.size());
  for (size_t t = 0; t < tris.size(); ++t) {
    const Triangle& T = tris[t];
    const int32_t v[6] = {T.pt1.x, T.pt1.y, T.pt2.x, T.pt2.y, T.pt3.x, T.pt3.y};
    std::memcpy(&tv[6 * t], v, sizeof(v));
  }
  check(hpm_prior_build(ctx_, tv.data(), (int)tris.size()), "hpm_prior_build");
}

void HPM::CudaPlanarPriorInitialization(const std::vector<float>& plane_params, const MapF& masks) {
  const size_t n = (size_t)masks.width * masks.height;
  std::vector<float> prior(4 * n, 0.0f);
  std::vector<uint32_t> m(n);
  for (size_t i = 0; i < n; ++i) {
    m[i] = (uint32_t)masks.data[i];
    if (masks.data[i] > 0) std::memcpy(&prior[4 * i], &plane_params[4 * (m[i] - 1)], 16);
  }
  check(hpm_set_prior(ctx_, prior.data(), m.data()), "hpm_set_prior");
}

void HPM::ReloadPlanarPriorInitialization(const MapF& masks, const std::vector<float>& prior_planes) {
  const size_t n = (size_t)masks.width * masks.height;
  std::vector<float> prior(4 * n, 0.0f);
  std::vector<uint32_t> m(n);
  for (size_t i = 0; i < n; ++i) {
    m[i] = (uint32_t)masks.data[i];
    if (masks.data[i] > 0) std::memcpy(&prior[4 * i], &prior_planes[4 * i], 16);
  }
  check(hpm_set_prior(ctx_, prior.data(), m.data()), "hpm_set_prior");
}

// Reference views of at least this many pixels run in the mixed-precision
// engine as a one-view fused launch with 16-column colour tiles and 4-tile
// XCD strips (hpm_batch); smaller ones, and the exact FP32 engine, in the
// context's own launch (32-column tiles).  Same results either way.
static std::atomic<int64_t> g_fused_min_pixels{4000000};

void SetFusedMinPixels(int64_t n) { g_fused_min_pixels = n; }

void HPM::RunPatchMatch(uint64_t seed) {
  have_results_ = false;
  ran_ = true;
  check(hpm_set_params(ctx_, &params_), "hpm_set_params");
  // At 3072 x 2048 with 10 sources the narrower tiles keep more of the source
  // gathers in the caches: 7-20 % less time per RunPatchMatch.
  const int64_t px = (int64_t)GetReferenceImageWidth() * GetReferenceImageHeight();
  if (params_.precision != HPM_PREC_FP32_EXACT && px >= g_fused_min_pixels) {
    // one batch per HPM object, made on the first large run and kept: its
    // stream, events and layout tables serve every later run (the tables are
    // rebuilt only when the context's buffers or sizes changed)
    if (!fused_) {
      if (hpm_batch_create(&fused_, &ctx_, 1) != HPM_OK) throw Error("hpm_batch_create failed");
      if (hpm_batch_set_shape(fused_, 16, 4) != HPM_OK)
        throw Error(std::string("hpm_batch_set_shape: ") + hpm_batch_last_error(fused_));
    }
    if (hpm_batch_run(fused_, &seed) != HPM_OK)
      throw Error(std::string("hpm_batch_run: ") + hpm_batch_last_error(fused_));
    return;  // hpm_batch_run returns once its stream has finished
  }
  check(hpm_run_patchmatch(ctx_, seed), "hpm_run_patchmatch");
  check(hpm_synchronize(ctx_), "hpm_synchronize");
}

float HPM::GetMinDepth() const { return cams_.at(0).depth_min * 0.6f; }  // params.depth_min, HPM.cpp:685
float HPM::GetMaxDepth() const { return cams_.at(0).depth_max * 1.2f; }  // params.depth_max, HPM.cpp:686

void HPM::GetPlanes(std::vector<float>& planes) const {
  planes.resize((size_t)4 * GetReferenceImageWidth() * GetReferenceImageHeight());
  check(hpm_get_planes(ctx_, planes.data()), "hpm_get_planes");
}
void HPM::GetDepthNormals(MapF& depths, MapF& normals) const {
  depths = MapF(GetReferenceImageWidth(), GetReferenceImageHeight());
  normals = MapF(GetReferenceImageWidth(), GetReferenceImageHeight(), 3);
  check(hpm_get_depth_normals(ctx_, depths.data.data(), normals.data.data()), "hpm_get_depth_normals");
}

void HPM::GetCosts(MapF& costs) const {
  costs = MapF(GetReferenceImageWidth(), GetReferenceImageHeight());
  check(hpm_get_costs(ctx_, costs.data.data()), "hpm_get_costs");
}
void HPM::GetTexture(MapF& texture) const {
  texture = MapF(GetReferenceImageWidth(), GetReferenceImageHeight());
  check(hpm_get_texture(ctx_, texture.data.data()), "hpm_get_texture");
}

// Get3DPointonRefCam, HPM.cpp:258-267.
static void PointOnRefCam(int x, int y, float depth, const hpm_camera& c, float X[3]) {
  X[0] = depth * ((float)x - c.K[2]) / c.K[0];
  X[1] = depth * ((float)y - c.K[5]) / c.K[4];
  X[2] = depth;
}

void HPM::GetPriorPlaneParams(const Triangle& t, const MapF& depths, float plane[4]) const {
  float X1[3], X2[3], X3[3];
  PointOnRefCam(t.pt1.x, t.pt1.y, depths.at(t.pt1.x, t.pt1.y), cams_[0], X1);
  PointOnRefCam(t.pt2.x, t.pt2.y, depths.at(t.pt2.x, t.pt2.y), cams_[0], X2);
  PointOnRefCam(t.pt3.x, t.pt3.y, depths.at(t.pt3.x, t.pt3.y), cams_[0], X3);
  PlaneThroughPoints(X1, X2, X3, plane);
}

// Get3DPointonRefCam_factor, HPM.cpp:247-256.
static void PointOnRefCamFactor(int x, int y, float depth, const hpm_camera& c, float factor, float X[3]) {
  X[0] = depth * ((float)x - c.K[2] * factor) / (c.K[0] * factor);
  X[1] = depth * ((float)y - c.K[5] * factor) / (c.K[4] * factor);
  X[2] = depth;
}

void HPM::GetPriorPlaneParamsFactor(const Triangle& t, const MapF& depths, float factor, float plane[4]) const {
  float X1[3], X2[3], X3[3];
  PointOnRefCamFactor(t.pt1.x, t.pt1.y, depths.at(t.pt1.x, t.pt1.y), cams_[0], factor, X1);
  PointOnRefCamFactor(t.pt2.x, t.pt2.y, depths.at(t.pt2.x, t.pt2.y), cams_[0], factor, X2);
  PointOnRefCamFactor(t.pt3.x, t.pt3.y, depths.at(t.pt3.x, t.pt3.y), cams_[0], factor, X3);
  PlaneThroughPoints(X1, X2, X3, plane);
}

float HPM::GetDepthFromPlaneParamFactor(const float pl[4], int x, int y, float factor) const {  // HPM.cpp:1174-1177
  const float* K = cams_[0].K;
  return -pl[3] * (K[0] * factor) /
         (((float)x - K[2] * factor) * pl[0] + (K[0] / K[4]) * ((float)y - K[5] * factor) * pl[1] + K[0] * factor * pl[2]);
}

float HPM::GetDepthFromPlaneParam(const float pl[4], int x, int y) const {  // HPM.cpp:1169-1172
  const float* K = cams_[0].K;
  return -pl[3] * K[0] / (((float)x - K[2]) * pl[0] + (K[0] / K[4]) * ((float)y - K[5]) * pl[1] + K[0] * pl[2]);
}

void HPM::TransformNormal(const float in[4], float out[4]) const {  // HPM.cpp:1179-1187
  const float* R = cams_[0].R;
  out[0] = R[0] * in[0] + R[3] * in[1] + R[6] * in[2];
  out[1] = R[1] * in[0] + R[4] * in[1] + R[7] * in[2];
  out[2] = R[2] * in[0] + R[5] * in[1] + R[8] * in[2];
  out[3] = in[3];
}

void HPM::TransformNormal2RefCam(const float in[4], float out[4]) const {  // HPM.cpp:1189-1197
  const float* R = cams_[0].R;
  out[0] = R[0] * in[0] + R[1] * in[1] + R[2] * in[2];
  out[1] = R[3] * in[0] + R[4] * in[1] + R[5] * in[2];
  out[2] = R[6] * in[0] + R[7] * in[1] + R[8] * in[2];
  out[3] = in[3];
}

float HPM::GetDistance2Origin(int x, int y, float depth, const float n[4]) const {  // HPM.cpp:1199-1206
  float X[3];
  PointOnRefCam(x, y, depth, cams_[0], X);
  return -(n[0] * X[0] + n[1] * X[1] + n[2] * X[2]);
}

void HPM::FetchResults() {
  if (have_results_) return;
  const size_t n = (size_t)GetReferenceImageWidth() * GetReferenceImageHeight();
  planes_host_.resize(4 * n);
  costs_host_.resize(n);
  texture_host_.resize(n);
  check(hpm_get_planes(ctx_, planes_host_.data()), "hpm_get_planes");
  check(hpm_get_costs(ctx_, costs_host_.data()), "hpm_get_costs");
  check(hpm_get_texture(ctx_, texture_host_.data()), "hpm_get_texture");
  have_results_ = true;
}

Float4 HPM::GetPlaneHypothesis(int index) {  // HPM.cpp:1001-1004
  FetchResults();
  const float* p = &planes_host_.at(4 * (size_t)index);
  return Float4{p[0], p[1], p[2], p[3]};
}

float HPM::GetCost(int index) {  // HPM.cpp:1006-1009
  FetchResults();
  return costs_host_.at((size_t)index);
}

float HPM::GetTexture(int index) {  // HPM.cpp:1011-1014
  FetchResults();
  return texture_host_.at((size_t)index);
}

void HPM::GetSupportPoints_Classify_Check(std::vector<Point2i>& support2DPoints, const MapF& costs,
                                          const MapF& confidences, const MapF& texture, float hpm_factor) {
  (void)texture;  // its flag is computed but never used (HPM.cpp:1041-1048)
  const int width = (int)((float)GetReferenceImageWidth() * hpm_factor);
  const int height = (int)((float)GetReferenceImageHeight() * hpm_factor);
  support2DPoints = GetSupportPoints(costs, confidences, width, height);
}

std::vector<Triangle> HPM::DelaunayTriangulation(const Rect& boundRC, const std::vector<Point2i>& points) {
  if (boundRC.x != 0 || boundRC.y != 0) throw Error("DelaunayTriangulation: the bounding rectangle starts at (0, 0)");
  return hpmhost::DelaunayTriangulation(boundRC.width, boundRC.height, points);
}

Float4 HPM::GetPriorPlaneParams(const Triangle& triangle, const MapF& depths) const {
  Float4 r;
  GetPriorPlaneParams(triangle, depths, &r.x);
  return r;
}

Float4 HPM::GetPriorPlaneParams_factor(const Triangle& triangle, const MapF& depths, float factor) const {
  Float4 r;
  GetPriorPlaneParamsFactor(triangle, depths, factor, &r.x);
  return r;
}

float HPM::GetDepthFromPlaneParam(const Float4& plane, int x, int y) const {
  return GetDepthFromPlaneParam(&plane.x, x, y);
}

float HPM::GetDepthFromPlaneParam_factor(const Float4& plane, int x, int y, float factor) const {
  return GetDepthFromPlaneParamFactor(&plane.x, x, y, factor);
}

Float4 HPM::TransformNormal(const Float4& plane) const {
  Float4 r;
  TransformNormal(&plane.x, &r.x);
  return r;
}

Float4 HPM::TransformNormal2RefCam(const Float4& plane) const {
  Float4 r;
  TransformNormal2RefCam(&plane.x, &r.x);
  return r;
}

float HPM::GetDistance2Origin(const Int2& p, float depth, const Float4& normal) const {
  return GetDistance2Origin(p.x, p.y, depth, &normal.x);
}

void HPM::CudaPlanarPriorInitialization(const std::vector<Float4>& plane_params, const MapF& masks) {
  std::vector<float> flat(4 * plane_params.size());
  if (!plane_params.empty()) std::memcpy(flat.data(), plane_params.data(), flat.size() * sizeof(float));
  CudaPlanarPriorInitialization(flat, masks);
}

void HPM::ReloadPlanarPriorInitialization(const MapF& masks, const Float4* prior_plane_parameters) {
  const size_t n = (size_t)masks.width * masks.height;
  std::vector<float> flat(4 * n);
  std::memcpy(flat.data(), prior_plane_parameters, flat.size() * sizeof(float));
  ReloadPlanarPriorInitialization(masks, flat);
}

void HPM::JointBilateralUpsampling_prior(const MapF& scaled_image_float, const MapF& src_depthmap,
                                         MapF& upsample_depthmap, const MapF& src_normal, MapF& upsample_normal) {
  const int w = scaled_image_float.width, h = scaled_image_float.height;
  const int cw = src_depthmap.width, ch = src_depthmap.height;
  if (src_normal.width != cw || src_normal.height != ch || src_normal.channels != 3)
    throw Error("JointBilateralUpsampling_prior: a 3-channel normal map of the depth map's size is needed");
  if (cw <= 0 || ch <= 0 || std::max(h / ch, w / cw) == 1) return;  // HPM.cpp:1327-1331: the outputs stay as they are
  std::vector<float> n4((size_t)4 * cw * ch, 0.0f);
  for (size_t i = 0; i < (size_t)cw * ch; ++i) std::memcpy(&n4[4 * i], &src_normal.data[3 * i], 3 * sizeof(float));
  upsample_depthmap = MapF(w, h);
  std::vector<float> up4((size_t)4 * w * h);
  if (hpm_jbu_prior(scaled_image_float.data.data(), w, h, src_depthmap.data.data(), n4.data(), cw, ch,
                    upsample_depthmap.data.data(), up4.data()) != HPM_OK)
    throw Error("hpm_jbu_prior failed");
  upsample_normal = MapF(w, h, 3);
  for (size_t i = 0; i < (size_t)w * h; ++i) std::memcpy(&upsample_normal.data[3 * i], &up4[4 * i], 3 * sizeof(float));
}

void HPM::CudaPlanarPriorRelease() { check(hpm_release_prior(ctx_), "hpm_release_prior"); }

void HPM::CudaSpaceRelease(bool geom_consistency) {
  (void)geom_consistency;  // the source depth maps go with the other buffers
  if (ran_) FetchResults();  // the getters keep answering from host copies
  ran_ = false;
  check(hpm_release(ctx_), "hpm_release");
}

void HPM::ReleaseProblemHostMemory() {
  images_.clear();
  owned_images_.clear();
  depths_.clear();
}

// -------------------------------------------------------------- pass driver

static void MakeDir(const std::string& path) { ::mkdir(path.c_str(), 0755); }

// Planar prior at distance 0 (main.cpp:326-441): support points, Delaunay,
// per-triangle planes, prior depth range test.
static void PriorAtLevel(HPM& hpm, const MapF& depths, const MapF& costs, const MapF& confidences) {
  const int width = hpm.GetReferenceImageWidth(), height = hpm.GetReferenceImageHeight();
  const std::vector<Point2i> support = GetSupportPoints(costs, confidences, width, height);
  const std::vector<Triangle> tris = DelaunayTriangulation(width, height, support);
  MapF mask(width, height);
  std::vector<float> planes;
  for (const Triangle& t : tris) {
    float pl[4];
    hpm.GetPriorPlaneParams(t, depths, pl);
    planes.insert(planes.end(), pl, pl + 4);
  }
  RasterTriangles(tris, mask);
  for (int i = 0; i < width; ++i)
    for (int j = 0; j < height; ++j) {
      if (mask.at(i, j) > 0) {
        const float d = hpm.GetDepthFromPlaneParam(&planes[4 * ((size_t)mask.at(i, j) - 1)], i, j);
        if (!(d <= hpm.GetMaxDepth() && d >= hpm.GetMinDepth())) mask.at(i, j) = 0;
      }
    }
  hpm.CudaPlanarPriorInitialization(planes, mask);
}

// Planar prior 1 or 2 levels coarser (main.cpp:443-597): the maps are
// downsampled by 1 / (2 d), the prior is built there, upsampled with the
// depth+normal JBU and reloaded per pixel.
static void PriorFromCoarser(HPM& hpm, const MapF& depths, const MapF& costs, const MapF& confidences, int distance) {
  const int width = hpm.GetReferenceImageWidth(), height = hpm.GetReferenceImageHeight();
  const float factor = (float)(1.0 / (distance * 2));
  const int hw = (int)std::round(width * factor), hh = (int)std::round(height * factor);
  const MapF depths_ds = ResizeLinear(depths, hw, hh);
  const MapF costs_ds = ResizeLinear(costs, hw, hh);
  const MapF conf_ds = ResizeLinear(confidences, hw, hh);
  // GetSupportPoints_Classify_Check scans (int)(width * factor) columns (HPM.cpp:1030-1031)
  const int sw = (int)(width * factor), sh = (int)(height * factor);
  const std::vector<Point2i> support = GetSupportPoints(costs_ds, conf_ds, std::min(sw, hw), std::min(sh, hh));
  const std::vector<Triangle> tris = DelaunayTriangulation(hw, hh, support);
  MapF mask(hw, hh);
  std::vector<float> planes;
  for (const Triangle& t : tris) {
    float pl[4];
    hpm.GetPriorPlaneParamsFactor(t, depths_ds, factor, pl);
    planes.insert(planes.end(), pl, pl + 4);
  }
  RasterTriangles(tris, mask);
  MapF prior_d(hw, hh);
  std::vector<float> prior_n((size_t)4 * hw * hh, 0.0f);  // world normal xyz, depth
  for (int i = 0; i < hw; ++i)
    for (int j = 0; j < hh; ++j) {
      if (mask.at(i, j) > 0) {
        const float* pl = &planes[4 * ((size_t)mask.at(i, j) - 1)];
        const float d = hpm.GetDepthFromPlaneParamFactor(pl, i, j, factor);
        if (d <= hpm.GetMaxDepth() * 1.2f && d >= hpm.GetMinDepth() * 0.6f) {
          prior_d.at(i, j) = d;
          float nw[4];
          hpm.TransformNormal(pl, nw);
          const size_t c = (size_t)j * hw + i;
          prior_n[4 * c + 0] = nw[0];
          prior_n[4 * c + 1] = nw[1];
          prior_n[4 * c + 2] = nw[2];
          prior_n[4 * c + 3] = d;
        } else {
          mask.at(i, j) = 0;
        }
      }
    }
  // JointBilateralUpsampling_prior (HPM.cpp:1323-1392): nothing is upsampled
  // when the integer scale is 1; its NaN patch is overwritten at once
  MapF up_d(width, height);
  std::vector<float> up_n((size_t)4 * width * height, 0.0f);
  const int scale = std::max(height / hh, width / hw);
  if (scale != 1) {
    const int rc = hpm_jbu_prior(hpm.GetReferenceImage().data.data(), width, height, prior_d.data.data(),
                                 prior_n.data(), hw, hh, up_d.data.data(), up_n.data());
    if (rc != HPM_OK) throw Error("hpm_jbu_prior failed");
  }
  MapF mask_new(width, height);
  std::vector<float> prior_planes((size_t)4 * width * height, 0.0f);
  for (int i = 0; i < width; ++i)
    for (int j = 0; j < height; ++j) {
      const float d = up_d.at(i, j);
      if (d <= hpm.GetMaxDepth() && d >= hpm.GetMinDepth()) {
        mask_new.at(i, j) = 1;
        const size_t c = (size_t)j * width + i;
        const float in[4] = {up_n[4 * c], up_n[4 * c + 1], up_n[4 * c + 2], d};
        float pl[4];
        hpm.TransformNormal2RefCam(in, pl);
        pl[3] = hpm.GetDistance2Origin(i, j, d, pl);
        std::memcpy(&prior_planes[4 * c], pl, 16);
      }
    }
  hpm.ReloadPlanarPriorInitialization(mask_new, prior_planes);
}

namespace {
std::mutex g_stage_mu;
StageTimes g_stages{};
int g_verbose = 0;

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Adds the time since construction (or the last lap) to a stage.
class StageClock {
 public:
  StageClock() : t_(now_s()) {}
  double lap(Stage s) {
    const double t = now_s(), d = t - t_;
    t_ = t;
    std::lock_guard<std::mutex> lk(g_stage_mu);
    g_stages.seconds[s] += d;
    g_stages.calls[s] += 1;
    return d;
  }
 private:
  double t_;
};
}  // namespace

StageTimes GetStageTimes() {
  std::lock_guard<std::mutex> lk(g_stage_mu);
  return g_stages;
}

void ResetStageTimes() {
  std::lock_guard<std::mutex> lk(g_stage_mu);
  g_stages = StageTimes{};
}

void SetVerbose(int level) { g_verbose = level; }

namespace {
int g_pass_threads = 0;  // 0: min(8, hardware threads)

int pass_threads(int n) {
  int t = g_pass_threads;
  if (t <= 0) t = std::min(8, (int)std::max(1u, std::thread::hardware_concurrency()));
  return std::max(1, std::min(t, n));
}

// Runs fn(i) for i in [0, n) on up to pass_threads(n) host threads; the first
// exception is rethrown after every thread has finished.
template <class F> void parallel_views(int n, F fn) {
  const int t = pass_threads(n);
  if (t <= 1) {
    for (int i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<int> next{0};
  std::exception_ptr err;
  std::mutex err_mu;
  std::vector<std::thread> pool;
  for (int k = 0; k < t; ++k)
    pool.emplace_back([&] {
      for (int i; (i = next.fetch_add(1)) < n;) {
        try {
          fn(i);
        } catch (...) {
          std::lock_guard<std::mutex> lk(err_mu);
          if (!err) err = std::current_exception();
          next.store(n);
        }
      }
    });
  for (auto& th : pool) th.join();
  if (err) std::rethrow_exception(err);
}

// Whether the views of a pass may run in any order: a pass reads, besides
// its own view's files, only files that no view of the pass writes.  Every
// pass qualifies except the geometric passes over the previous pass's
// depths_geom.dmb (multi_geometry), which the reference runs in view order,
// each view reading the sources already updated before it.
bool order_free(const PassFlags& f) { return !(f.geom_consistency && f.multi_geometry); }
}  // namespace

void SetPassThreads(int n) { g_pass_threads = n; }

namespace {
std::atomic<int> g_prior_on_device{1};
}
void SetPriorOnDevice(int enable) { g_prior_on_device = enable != 0; }

namespace {
std::atomic<int> g_confidence_overlap{1};

// ConfidenceEvaluation, then `pass(ready)`.  With the overlap on, the pass
// runs alongside the evaluation (a host pass with little GPU work) and each
// of its views starts once its own confidence map -- the only one it reads
// -- is written, i.e. once the evaluation has finished that image; the files
// are those of running one after the other.
template <class Pass>
void ConfidenceThenPass(const std::string& dense_folder, const std::vector<Problem>& problems, bool geom, Pass pass) {
  if (!g_confidence_overlap) {
    ConfidenceEvaluation(dense_folder, problems, geom);
    pass(nullptr);
    return;
  }
  ReadyViews ready((int)problems.size());
  std::exception_ptr conf_err;
  std::thread conf([&] {
    try {
      ConfidenceEvaluation(dense_folder, problems, geom, &ready);
    } catch (...) {
      conf_err = std::current_exception();
      ready.Fail("confidence evaluation failed");
    }
  });
  try {
    pass(&ready);
  } catch (...) {
    conf.join();
    if (conf_err) std::rethrow_exception(conf_err);
    throw;
  }
  conf.join();
  if (conf_err) std::rethrow_exception(conf_err);
}
}  // namespace

void SetConfidenceOverlap(int enable) { g_confidence_overlap = enable != 0; }

// A PatchMatch pass takes the reference and 1 to HPM_MAX_IMAGES - 1 source
// views, a repeated source counting each time.  main.cpp:1209-1220 refuses
// a problem over MAX_IMAGES before any pass; a view whose sources all have
// score <= 0 in pair.txt is refused here as well, since its costs would be
// averages over no view.
static void CheckSources(const Problem& p) {
  if (p.src_image_ids.empty())
    throw Error("view " + std::to_string(p.ref_image_id) + " has no source view with a positive score in pair.txt");
  if (1 + p.src_image_ids.size() > (size_t)HPM_MAX_IMAGES)
    throw Error("view " + std::to_string(p.ref_image_id) + " needs " + std::to_string(1 + p.src_image_ids.size()) +
                " images, more than HPM_MAX_IMAGES = " + std::to_string(HPM_MAX_IMAGES));
}

static void CheckSources(const std::vector<Problem>& problems) {
  for (const Problem& p : problems) CheckSources(p);
}

void ProcessProblem(const std::string& dense_folder, const std::vector<Problem>& problems, int idx,
                    const PassFlags& flags, int precision, uint64_t seed, SceneImages* images) {
  StageClock clk;
  double t_setup = 0, t_edges = 0, t_prior = 0, t_pm = 0, t_write = 0;
  const Problem& problem = problems.at(idx);
  CheckSources(problem);
  MakeDir(dense_folder + "/HPM_MVS_plusplus");
  const std::string rf = ResultFolder(dense_folder, problem.ref_image_id);
  MakeDir(rf);
  HPM hpm;
  hpm.SetPrecision(precision);
  if (flags.geom_consistency) hpm.SetGeomConsistencyParams(flags.multi_geometry);
  if (flags.hierarchy) hpm.SetHierarchyParams();
  hpm.InuputInitialization(dense_folder, problems, idx, images);
  hpm.SetMandConsistencyParams(flags.mand_consistency);
  if (flags.prior_consistency) hpm.SetPlanarPriorParams();
  hpm.CudaSpaceInitialization(dense_folder, problem);
  const int width = hpm.GetReferenceImageWidth(), height = hpm.GetReferenceImageHeight();
  t_setup += clk.lap(kSetup);

  // cv::Canny on the level's colour image, main.cpp:280-289
  if (images) {
    const MapU8& edges = images->CurrentEdges(problem.ref_image_id);
    if (edges.width != width || edges.height != height)
      throw Error("prepared reference image size does not match the camera size");
    hpm.CudaCannyInitialization(edges);
  } else {
    const MapU8 colour = LoadImage(dense_folder, problem.ref_image_id);
    if (colour.width != width || colour.height != height)
      throw Error("reference image size does not match the camera size");
    hpm.CudaCannyInitialization(DeviceEdges(colour));
  }
  t_edges += clk.lap(kEdges);
  if (flags.mand_consistency || flags.prior_consistency) hpm.CudaConfidenceInitialization(dense_folder, problems, idx);

  bool run = true;
  MapRef in_depths, in_normals, in_costs;  // a prior pass's loaded maps
  if (flags.prior_consistency) {  // main.cpp:324-597
    const char* suffix = flags.multi_geometry ? "/depths_geom.dmb" : "/depths.dmb";
    in_depths = LoadMap(rf + suffix);
    in_normals = LoadMap(rf + "/normals.dmb");
    const MapRef conf_ref = LoadMap(rf + "/confidence.dmb");
    in_costs = LoadMap(rf + "/costs.dmb");
    const MapF& depths = Shaped(in_depths, width, height, 1, rf + suffix);
    const MapF& normals = Shaped(in_normals, width, height, 3, rf + "/normals.dmb");
    const MapF& costs = Shaped(in_costs, width, height, 1, rf + "/costs.dmb");
    const MapF& confidences = Shaped(conf_ref, width, height, 1, rf + "/confidence.dmb");
    // the prior texture map must exist (main.cpp:353-355, 458-460); its
    // texture flag is never used by the support-point selection
    const int tex_level = flags.image_scale + flags.hpm_scale_distance;
    if (flags.hpm_scale_distance >= 0 && flags.hpm_scale_distance <= 2)
      (void)LoadMap(rf + "/texture" + std::to_string(tex_level) + ".dmb");
    t_setup += clk.lap(kSetup);
    // any distance but 0, 1, 2 runs nothing: the loaded maps are written back
    // unchanged, as main.cpp:349-597 does when neither branch is taken
    run = flags.hpm_scale_distance >= 0 && flags.hpm_scale_distance <= 2;
    if (run) {
      // the reload and the prior fill separate buffers; the device build
      // reads the reloaded depths and costs
      hpm.CudaHypothesesReload(depths, costs, normals);
      if (g_prior_on_device)
        hpm.BuildPlanarPrior(flags.hpm_scale_distance);
      else if (flags.hpm_scale_distance == 0)
        PriorAtLevel(hpm, depths, costs, confidences);
      else
        PriorFromCoarser(hpm, depths, costs, confidences, flags.hpm_scale_distance);
    }
    t_prior += clk.lap(kPrior);
  } else {
    t_setup += clk.lap(kSetup);
  }
  MapF depths, normals, costs;
  double t_run = 0;
  if (run) {
    const double r0 = now_s();
    hpm.RunPatchMatch(seed);
    t_run = now_s() - r0;
    hpm.GetDepthNormals(depths, normals);  // main.cpp:311-322
    hpm.GetCosts(costs);
    t_pm += clk.lap(kPatchMatch);
  } else {  // the loaded maps go back unchanged
    depths = *in_depths;
    normals = *in_normals;
    costs = *in_costs;
  }
  const std::string dpath = rf + (flags.geom_consistency ? "/depths_geom.dmb" : "/depths.dmb");
  SaveMap(dpath, std::move(depths));
  SaveMap(rf + "/normals.dmb", std::move(normals));
  SaveMap(rf + "/costs.dmb", std::move(costs));
  if (!flags.mand_consistency && !flags.prior_consistency) {
    MapF texture;
    hpm.GetTexture(texture);
    SaveMap(rf + "/texture" + std::to_string(flags.image_scale) + ".dmb", std::move(texture));
  }
  t_write += clk.lap(kWrite);
  if (g_verbose > 0) {
    const char* mode = flags.prior_consistency ? "prior"
                       : flags.hierarchy       ? "hierarchy"
                       : flags.mand_consistency ? "geometric+mandatory"
                       : flags.geom_consistency ? "geometric"
                                                : "photometric";
    std::fprintf(stderr,
                 "[hpm] view %08d %dx%d %s%s: setup %.1f ms, edges %.1f, prior %.1f, patchmatch %.1f (run %.1f, "
                 "read back %.1f), write %.1f\n",
                 problem.ref_image_id, width, height, mode,
                 flags.prior_consistency ? (" d=" + std::to_string(flags.hpm_scale_distance)).c_str() : "",
                 1e3 * t_setup, 1e3 * t_edges, 1e3 * t_prior, 1e3 * t_pm, 1e3 * t_run, 1e3 * (t_pm - t_run),
                 1e3 * t_write);
    std::fflush(stderr);
  }
}

void JointBilateralUpsampling(const std::string& dense_folder, const Problem& problem, SceneImages& images) {
  StageClock clk;
  struct Lap {
    StageClock& c;
    ~Lap() { c.lap(kJbu); }
  } lap{clk};
  const std::string rf = ResultFolder(dense_folder, problem.ref_image_id);
  const MapRef coarse_ref = LoadMap(rf + "/depths_geom.dmb");
  const MapF& coarse = Shaped(coarse_ref, coarse_ref->width, coarse_ref->height, 1, rf + "/depths_geom.dmb");
  const MapF& fine = images.CurrentGray(problem.ref_image_id);
  const int scale = std::max(fine.height / coarse.height, fine.width / coarse.width);  // HPM.cpp:1270
  if (scale == 1) return;  // HPM.cpp:1272-1275
  MapF out(fine.width, fine.height);
  if (hpm_jbu_depth(fine.data.data(), fine.width, fine.height, coarse.data.data(), coarse.width, coarse.height,
                    out.data.data()) != HPM_OK)
    throw Error("hpm_jbu_depth failed");
  for (int i = 0; i < fine.width; ++i)  // NaN patch, HPM.cpp:1297-1300
    for (int j = 0; j < fine.height; ++j)
      if (out.at(i, j) != out.at(i, j))
        out.at(i, j) = coarse.at(std::min(i / 2, coarse.width - 1), std::min(j / 2, coarse.height - 1));
  SaveMap(rf + "/depths.dmb", std::move(out));
}

struct ViewMaps {
  std::vector<hpm_camera> cams;
  std::vector<MapRef> depths, normals;
  std::vector<int32_t> offsets, ids;
};

static ViewMaps LoadViewMaps(const std::string& dense_folder, const std::vector<Problem>& problems, bool geom) {
  // The cameras are rescaled to the depth-map size (RescaleImageAndCamera,
  // HPM.cpp:186-207); here the maps have the image size.
  ViewMaps v;
  for (const Problem& p : problems) {
    const std::string rf = ResultFolder(dense_folder, p.ref_image_id);
    const std::string dpath = rf + (geom ? "/depths_geom.dmb" : "/depths.dmb");
    v.depths.push_back(LoadMap(dpath));
    const MapF& dm = Shaped(v.depths.back(), v.depths.back()->width, v.depths.back()->height, 1, dpath);
    v.normals.push_back(LoadMap(rf + "/normals.dmb"));
    Shaped(v.normals.back(), dm.width, dm.height, 3, rf + "/normals.dmb");
    std::ostringstream cp;
    cp << dense_folder << "/cams/" << std::setw(8) << std::setfill('0') << p.ref_image_id << "_cam.txt";
    hpm_camera c = ReadCamera(cp.str());
    int iw = 0, ih = 0;
    ImageSize(dense_folder, p.ref_image_id, iw, ih);
    const MapF& d = *v.depths.back();
    if (iw != d.width || ih != d.height) {
      const float sx = d.width / (float)iw, sy = d.height / (float)ih;
      c.K[0] *= sx; c.K[2] *= sx; c.K[4] *= sy; c.K[5] *= sy;
    }
    c.width = d.width;
    c.height = d.height;
    v.cams.push_back(c);
    v.offsets.push_back(0);  // filled below
  }
  // main.cpp:897 indexes the loaded maps with the source image id; map ids to
  // positions in pair.txt order
  std::vector<int> pos;
  for (size_t i = 0; i < problems.size(); ++i) {
    const int id = problems[i].ref_image_id;
    if (id < 0) throw Error("negative image id in pair.txt");
    if ((size_t)id >= pos.size()) pos.resize(id + 1, -1);
    pos[id] = (int)i;
  }
  v.offsets.assign(1, 0);
  for (const Problem& p : problems) {
    for (int s : p.src_image_ids) {
      if (s < 0 || (size_t)s >= pos.size() || pos[s] < 0) throw Error("source image without a pair.txt entry");
      v.ids.push_back(pos[s]);
    }
    v.offsets.push_back((int32_t)v.ids.size());
  }
  return v;
}

ReadyViews::ReadyViews(int n) : ready_((size_t)n, 0) {}

void ReadyViews::Mark(int i) {
  {
    std::lock_guard<std::mutex> l(mu_);
    ready_.at(i) = 1;
  }
  cv_.notify_all();
}

void ReadyViews::Fail(const std::string& msg) {
  {
    std::lock_guard<std::mutex> l(mu_);
    if (error_.empty()) error_ = msg;
  }
  cv_.notify_all();
}

void ReadyViews::Wait(int i) {
  std::unique_lock<std::mutex> l(mu_);
  cv_.wait(l, [&] { return ready_.at(i) || !error_.empty(); });
  if (!ready_[i]) throw Error(error_);
}

void ConfidenceEvaluation(const std::string& dense_folder, const std::vector<Problem>& problems, bool geom,
                          ReadyViews* ready) {
  StageClock clk;
  struct Lap {
    StageClock& c;
    ~Lap() { c.lap(kConfidence); }
  } lap{clk};
  const double t0 = now_s();
  ViewMaps v = LoadViewMaps(dense_folder, problems, geom);
  const double t1 = now_s();
  const int n = (int)problems.size();
  std::vector<MapF> conf;
  std::vector<const float*> dp, np;
  std::vector<float*> cp;
  for (int i = 0; i < n; ++i) {
    conf.emplace_back(v.depths[i]->width, v.depths[i]->height);
    dp.push_back(v.depths[i]->data.data());
    np.push_back(v.normals[i]->data.data());
  }
  for (auto& c : conf) cp.push_back(c.data.data());
  // The reference writes image i's confidence.dmb as soon as image i is done
  // (main.cpp:1095-1100): the file holds what images 0..i wrote into the map.
  // hpm_consistency_check_cb fills conf[i] with exactly that when image i is
  // done, reports it, and never writes conf[i] again (later images' writes
  // go to its own in-memory maps), so the map is saved as it is.
  struct Progress {
    const std::string* folder;
    const std::vector<Problem>* problems;
    std::vector<MapF>* conf;
    ReadyViews* ready;
  } prog{&dense_folder, &problems, &conf, ready};
  auto image_done = [](int k, void* user) {
    auto* g = static_cast<Progress*>(user);
    SaveMap(ResultFolder(*g->folder, (*g->problems)[k].ref_image_id) + "/confidence.dmb", std::move((*g->conf)[k]));
    if (g->ready) g->ready->Mark(k);
  };
  int64_t count = 0;
  const int rc = hpm_consistency_check_cb(HPM_CHECK_CONFIDENCE, n, v.cams.data(), dp.data(), np.data(),
                                          v.offsets.data(), v.ids.data(), cp.data(), nullptr, nullptr, 0, &count,
                                          image_done, &prog);
  if (rc != HPM_OK) throw Error("hpm_consistency_check (confidence) failed");
  const double t2 = now_s();
  if (g_verbose > 0) {
    std::fprintf(stderr, "[hpm] confidence (%d views): load %.1f ms, check and hand-over %.1f ms\n", n,
                 1e3 * (t1 - t0), 1e3 * (t2 - t1));
    std::fflush(stderr);
  }
}

namespace {
// The fused points of RunFusion's consistency pass (main.cpp:823-957), with
// the view and pixel each one comes from.
struct Fused {
  std::vector<FusedPoint> pts;
  std::vector<int32_t> view, pixel;
  std::vector<int> width, height;  // per view: the depth map size
};

Fused FuseViews(const std::string& dense_folder, const std::vector<Problem>& problems, bool geom, int mode) {
  ViewMaps v = LoadViewMaps(dense_folder, problems, geom);
  const int n = (int)problems.size();
  std::vector<const float*> dp, np;
  size_t cap = 0;
  for (int i = 0; i < n; ++i) {
    dp.push_back(v.depths[i]->data.data());
    np.push_back(v.normals[i]->data.data());
    cap += v.depths[i]->data.size();
  }
  // room for every pixel, left uninitialised: only the fused ones are written
  std::unique_ptr<int32_t[]> pairs(new int32_t[2 * std::max<size_t>(cap, 1)]);
  std::unique_ptr<float[]> xyzn(new float[6 * std::max<size_t>(cap, 1)]);
  int64_t count = 0;
  const int rc = hpm_consistency_check(mode, n, v.cams.data(), dp.data(), np.data(), v.offsets.data(),
                                       v.ids.data(), nullptr, pairs.get(), xyzn.get(), (int64_t)cap, &count);
  if (rc != HPM_OK) throw Error("hpm_consistency_check (fusion) failed");
  // colours from the image resized to its depth map (RescaleImageAndCamera,
  // HPM.cpp:186-207)
  std::vector<MapU8> imgs;
  Fused F;
  for (int i = 0; i < n; ++i) {
    MapU8 img = LoadImage(dense_folder, problems[i].ref_image_id);
    if (img.width != v.depths[i]->width || img.height != v.depths[i]->height)
      img = ResizeLinear(img, v.depths[i]->width, v.depths[i]->height);
    imgs.push_back(std::move(img));
    F.width.push_back(v.depths[i]->width);
    F.height.push_back(v.depths[i]->height);
  }
  F.pts.resize((size_t)count);
  F.view.resize((size_t)count);
  F.pixel.resize((size_t)count);
  for (int64_t k = 0; k < count; ++k) {
    FusedPoint& fp = F.pts[k];
    std::memcpy(fp.coord, &xyzn[6 * k], 12);
    std::memcpy(fp.normal, &xyzn[6 * k + 3], 12);
    const MapU8& img = imgs[pairs[2 * k]];
    const size_t p = (size_t)pairs[2 * k + 1];
    for (int c = 0; c < 3; ++c)  // B, G, R (a grey image repeats its value)
      fp.color[c] = img.channels == 1 ? img.data[p] : img.data[3 * p + (2 - c)];
    F.view[k] = pairs[2 * k];
    F.pixel[k] = pairs[2 * k + 1];
  }
  return F;
}

void ExportBoth(const std::string& dense_folder, const char* name, const std::vector<FusedPoint>& pts) {
  const std::string out = dense_folder + "/HPM_MVS_plusplus";
  MakeDir(out);
  bool fused_ok = false;
  std::thread fused([&] { fused_ok = ExportFusedPointCloud(out + "/fused.ply", pts); });  // the two files at once
  const bool ok = ExportPointCloud(out + "/" + name, pts);
  fused.join();
  if (!ok || !fused_ok) throw Error("cannot write the point clouds under " + out);
}
}  // namespace

std::vector<FusedPoint> RunFusion(const std::string& dense_folder, const std::vector<Problem>& problems, bool geom) {
  StageClock clk;
  struct Lap {
    StageClock& c;
    ~Lap() { c.lap(kFusion); }
  } lap{clk};
  Fused F = FuseViews(dense_folder, problems, geom, HPM_CHECK_FUSION);
  ExportBoth(dense_folder, "HPM_MVS_plusplus.ply", F.pts);
  return std::move(F.pts);
}

std::vector<FusedPoint> RunFusionSkyStrict(const std::string& dense_folder, const std::vector<Problem>& problems,
                                           bool geom) {
  StageClock clk;
  struct Lap {
    StageClock& c;
    ~Lap() { c.lap(kFusion); }
  } lap{clk};
  Fused F = FuseViews(dense_folder, problems, geom, HPM_CHECK_FUSION_MASK);
  // masks/%08d resized to the depth map (RescaleMask, HPM.cpp:209-219), B G R
  const int n = (int)problems.size();
  std::vector<MapU8> masks;
  for (int i = 0; i < n; ++i) {
    MapU8 m;
    const std::string base = dense_folder + "/masks/" + [&] {
      std::ostringstream os;
      os << std::setw(8) << std::setfill('0') << problems[i].ref_image_id;
      return os.str();
    }();
    if (!ReadNetpbm(base + ".ppm", m) && !ReadNetpbm(base + ".pgm", m))
      throw Error("cannot read mask " + base + ".ppm (binary PPM/PGM)");
    if (m.width != F.width[i] || m.height != F.height[i]) m = ResizeLinear(m, F.width[i], F.height[i]);
    masks.push_back(std::move(m));
  }
  std::vector<FusedPoint> kept;
  kept.reserve(F.pts.size());
  for (size_t k = 0; k < F.pts.size(); ++k) {
    const MapU8& m = masks[F.view[k]];
    const size_t p = (size_t)F.pixel[k];
    const int b = m.channels == 1 ? m.data[p] : m.data[3 * p + 2];
    const int g = m.channels == 1 ? m.data[p] : m.data[3 * p + 1];
    const int r = m.channels == 1 ? m.data[p] : m.data[3 * p + 0];
    // main.cpp:805: any channel at its sky value rejects the point
    if (b != 234 && g != 235 && r != 55) kept.push_back(F.pts[k]);
  }
  ExportBoth(dense_folder, "HPM_MVS_plusplus_mask.ply", kept);
  return kept;
}

std::vector<FusedPoint> RunPipeline(const std::string& dense_folder, int precision, uint64_t seed) {
  MapStoreScope store;  // the passes' maps stay in memory for the passes after them
  const std::vector<Problem> problems = ReadPairFile(dense_folder + "/pair.txt");
  CheckSources(problems);
  const int n = (int)problems.size();
  PassFlags f;
  parallel_views(n, [&](int i) { ProcessProblem(dense_folder, problems, i, f, precision, seed + i); });  // photometric
  f.prior_consistency = true;
  ConfidenceThenPass(dense_folder, problems, false, [&](ReadyViews* ready) {
    parallel_views(n, [&](int i) {
      if (ready) ready->Wait(i);
      ProcessProblem(dense_folder, problems, i, f, precision, seed + 1000 + i);
    });
  });
  f = PassFlags();
  for (int it = 0; it < 3; ++it) {  // main.cpp:1269-1290
    f.multi_geometry = it > 0;
    const bool conf_geom = f.geom_consistency;  // the previous pass's maps
    if (it > 0) f.mand_consistency = true;
    f.geom_consistency = true;
    auto run = [&](ReadyViews* ready) {
      auto one = [&](int i) {
        if (ready) ready->Wait(i);
        ProcessProblem(dense_folder, problems, i, f, precision, seed + 2000 * (it + 2) + i);
      };
      if (order_free(f)) {
        parallel_views(n, one);
      } else {
        for (int i = 0; i < n; ++i) one(i);
      }
    };
    if (it > 0) ConfidenceThenPass(dense_folder, problems, conf_geom, run);
    else run(nullptr);
  }
  std::vector<FusedPoint> pts = RunFusion(dense_folder, problems, true);
  store.Finish();
  return pts;
}

std::vector<FusedPoint> RunPyramidPipeline(const std::string& dense_folder, int initial_image_scale, int size_bound,
                                           int precision, uint64_t seed, bool mask_fusion) {
  if (initial_image_scale < 1) throw Error("initial image scale must be >= 1");
  if (size_bound < 1) throw Error("size bound must be >= 1");
  MapStoreScope store;  // the passes' maps stay in memory for the passes after them
  const std::vector<Problem> problems = ReadPairFile(dense_folder + "/pair.txt");
  CheckSources(problems);
  const int n = (int)problems.size();
  MakeDir(dense_folder + "/HPM_MVS_plusplus");
  SceneImages images(dense_folder);
  // ComputeMultiScaleSettings, main.cpp:202-235
  std::vector<int> initial_max(n), num_down(n), cur_max(n, 0);
  int max_down = -1;
  for (int i = 0; i < n; ++i) {
    const MapU8& img = images.Original(problems[i].ref_image_id);
    const int side = std::max(img.width, img.height);
    initial_max[i] = std::max(1, (side + initial_image_scale / 2) / initial_image_scale);
    int k = 0;
    for (int m = initial_max[i]; m > size_bound; m /= 2) ++k;
    num_down[i] = k;
    max_down = std::max(max_down, k);
  }
  // every image a problem reads takes the level size of the problem whose
  // reference it is (PyramidMaxSideForImage, main.cpp:93-99)
  std::vector<int> ids;
  for (const Problem& p : problems) {
    ids.push_back(p.ref_image_id);
    ids.insert(ids.end(), p.src_image_ids.begin(), p.src_image_ids.end());
  }
  std::sort(ids.begin(), ids.end());
  ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
  auto level_side = [&](int id) {
    for (int i = 0; i < n; ++i)
      if (problems[i].ref_image_id == id) return cur_max[i];
    throw Error("no pyramid size for source image " + std::to_string(id));
  };

  uint64_t pass = 0;
  // one pass over the views; with `ready`, each view first waits for its
  // confidence map, and an order-free pass takes the views in the order the
  // maps become ready
  auto each_view = [&](const PassFlags& f, const std::vector<int>* distances, ReadyViews* ready) {
    auto one = [&](int i) {
      if (ready) ready->Wait(i);
      PassFlags fi = f;
      if (distances) fi.hpm_scale_distance = (*distances)[i];
      ProcessProblem(dense_folder, problems, i, fi, precision, seed + 10007 * pass + i, &images);
    };
    const double t0 = now_s();
    const bool concurrent = order_free(f);
    if (concurrent) {
      parallel_views(n, one);
    } else {
      for (int i = 0; i < n; ++i) one(i);
    }
    if (g_verbose > 0) {
      std::fprintf(stderr, "[hpm] level %d pass %llu (%s, %s): %d views in %.1f ms\n", f.image_scale,
                   (unsigned long long)pass,
                   f.prior_consistency ? "prior" : f.hierarchy ? "hierarchy" : f.geom_consistency ? "geometric" : "photometric",
                   concurrent ? "concurrent" : "in view order", n, 1e3 * (now_s() - t0));
      std::fflush(stderr);
    }
    ++pass;
  };
  const int max_hpm_scale = max_down;
  for (int level = max_down; level >= 0; --level) {  // main.cpp:1235-1352
    for (int i = 0; i < n; ++i)
      if (num_down[i] >= 0) {
        cur_max[i] = (int)(initial_max[i] / std::pow(2.0, num_down[i]));
        --num_down[i];
      }
    for (int id : ids) images.PrepareLevel(id, level_side(id));
    images.BuildPool(ids);

    PassFlags f;
    f.image_scale = level;
    if (level != max_down) {  // upsample the coarser level, then a hierarchy pass
      parallel_views(n, [&](int i) { JointBilateralUpsampling(dense_folder, problems[i], images); });
      f.hierarchy = true;
    }
    each_view(f, nullptr, nullptr);  // photometric or hierarchy pass
    f.hierarchy = false;
    f.prior_consistency = true;
    for (int hs = max_hpm_scale; hs >= level; --hs) {  // prior passes, coarsest prior first
      std::vector<int> dist(n);
      for (int i = 0; i < n; ++i) dist[i] = hs - num_down[i] - 1;
      ConfidenceThenPass(dense_folder, problems, false, [&](ReadyViews* ready) { each_view(f, &dist, ready); });
    }
    f.prior_consistency = false;
    f.hpm_scale_distance = 0;
    for (int it = 0; it < 3; ++it) {  // geometric passes
      f.multi_geometry = it > 0;
      f.mand_consistency = it > 0;
      f.geom_consistency = true;
      if (it > 0)  // over the previous depths_geom
        ConfidenceThenPass(dense_folder, problems, true, [&](ReadyViews* ready) { each_view(f, nullptr, ready); });
      else
        each_view(f, nullptr, nullptr);
    }
  }
  std::vector<FusedPoint> pts =
      mask_fusion ? RunFusionSkyStrict(dense_folder, problems, true) : RunFusion(dense_folder, problems, true);
  store.Finish();
  return pts;
}

}  // namespace hpmhost
