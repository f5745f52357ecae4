// ba_vimap.cc -- maplab map folders to the visual BA problem and back
// (include/maplab_ba_vimap.h): a protobuf wire-format reader and writer for
// the VIMap messages the BA needs (field numbers of vi_map.proto,
// visual-frame.proto, id.proto), gzip through zlib, and the subset of YAML
// that sensors.yaml uses (block mappings and sequences, flow sequences of
// scalars). The conversion to the problem follows
// OptimizationStateBuffer::importStatesOfMissions
// (algorithms/map-optimization/src/optimization-state-buffer.cc:153-263).
#include "../../include/maplab_ba_vimap.h"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

using Key = std::pair<uint64_t, uint64_t>;

// ------------------------------------------------------------ files, gzip

bool read_file(const std::string& path, std::string* out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  *out = ss.str();
  return true;
}

bool write_file(const std::string& path, const std::string& data) {
  const std::string tmp = path + ".tmp";
  {
    std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
    if (!f) return false;
    f.write(data.data(), (std::streamsize)data.size());
    if (!f) return false;
  }
  return std::rename(tmp.c_str(), path.c_str()) == 0;
}

bool gunzip(const std::string& in, std::string* out) {
  z_stream zs{};
  if (inflateInit2(&zs, 16 + MAX_WBITS) != Z_OK) return false;
  zs.next_in = reinterpret_cast<Bytef*>(const_cast<char*>(in.data()));
  zs.avail_in = (uInt)in.size();
  out->clear();
  char buf[1 << 16];
  int rc;
  do {
    zs.next_out = reinterpret_cast<Bytef*>(buf);
    zs.avail_out = sizeof(buf);
    rc = inflate(&zs, Z_NO_FLUSH);
    if (rc != Z_OK && rc != Z_STREAM_END) { inflateEnd(&zs); return false; }
    out->append(buf, sizeof(buf) - zs.avail_out);
  } while (rc != Z_STREAM_END);
  inflateEnd(&zs);
  return true;
}

bool gzip(const std::string& in, std::string* out) {
  z_stream zs{};
  if (deflateInit2(&zs, Z_DEFAULT_COMPRESSION, Z_DEFLATED, 16 + MAX_WBITS, 8, Z_DEFAULT_STRATEGY) != Z_OK)
    return false;
  zs.next_in = reinterpret_cast<Bytef*>(const_cast<char*>(in.data()));
  zs.avail_in = (uInt)in.size();
  out->clear();
  char buf[1 << 16];
  int rc;
  do {
    zs.next_out = reinterpret_cast<Bytef*>(buf);
    zs.avail_out = sizeof(buf);
    rc = deflate(&zs, Z_FINISH);
    if (rc == Z_STREAM_ERROR) { deflateEnd(&zs); return false; }
    out->append(buf, sizeof(buf) - zs.avail_out);
  } while (rc != Z_STREAM_END);
  deflateEnd(&zs);
  return true;
}

// ------------------------------------------------------------ protobuf wire format

struct Span {
  const uint8_t* p = nullptr;
  size_t n = 0;
};

struct Field {
  int num = 0, wt = 0;
  uint64_t v = 0;     // varint / fixed64 / fixed32 value
  Span s;             // length-delimited payload
  Span raw;           // the whole field, tag included
};

bool read_varint(const uint8_t*& p, const uint8_t* end, uint64_t* v) {
  uint64_t x = 0;
  for (int shift = 0; shift < 64 && p < end; shift += 7) {
    const uint8_t b = *p++;
    x |= (uint64_t)(b & 0x7F) << shift;
    if (!(b & 0x80)) { *v = x; return true; }
  }
  return false;
}

// Iterates the fields of one message; ok() is false after a malformed field.
class Reader {
 public:
  explicit Reader(Span s) : p_(s.p), end_(s.p + s.n) {}
  bool next(Field* f) {
    if (p_ >= end_) return false;
    const uint8_t* start = p_;
    uint64_t tag;
    if (!read_varint(p_, end_, &tag)) return bad();
    f->num = (int)(tag >> 3);
    f->wt = (int)(tag & 7);
    switch (f->wt) {
      case 0:
        if (!read_varint(p_, end_, &f->v)) return bad();
        break;
      case 1:
        if (end_ - p_ < 8) return bad();
        std::memcpy(&f->v, p_, 8);
        p_ += 8;
        break;
      case 2: {
        uint64_t len;
        if (!read_varint(p_, end_, &len) || (uint64_t)(end_ - p_) < len) return bad();
        f->s = Span{p_, (size_t)len};
        p_ += len;
        break;
      }
      case 5: {
        if (end_ - p_ < 4) return bad();
        uint32_t w;
        std::memcpy(&w, p_, 4);
        f->v = w;
        p_ += 4;
        break;
      }
      default:
        return bad();
    }
    f->raw = Span{start, (size_t)(p_ - start)};
    return true;
  }
  bool ok() const { return ok_; }

 private:
  bool bad() {
    ok_ = false;
    p_ = end_;
    return false;
  }
  const uint8_t* p_;
  const uint8_t* end_;
  bool ok_ = true;
};

// Repeated double (unpacked fixed64 fields or one packed run).
void add_doubles(const Field& f, std::vector<double>* out) {
  if (f.wt == 1) {
    double d;
    std::memcpy(&d, &f.v, 8);
    out->push_back(d);
  } else if (f.wt == 2) {
    for (size_t k = 0; k + 8 <= f.s.n; k += 8) {
      double d;
      std::memcpy(&d, f.s.p + k, 8);
      out->push_back(d);
    }
  }
}

// aslam.proto.Id: repeated uint64 uint = 1 -> (u[0], u[1]), (0, 0) otherwise.
Key parse_id(Span s) {
  std::vector<uint64_t> u;
  Reader r(s);
  Field f;
  while (r.next(&f)) {
    if (f.num != 1) continue;
    if (f.wt == 0) {
      u.push_back(f.v);
    } else if (f.wt == 2) {
      const uint8_t* p = f.s.p;
      const uint8_t* e = f.s.p + f.s.n;
      uint64_t v;
      while (p < e && read_varint(p, e, &v)) u.push_back(v);
    }
  }
  return u.size() == 2 ? Key(u[0], u[1]) : Key(0, 0);
}

void put_varint(std::string* out, uint64_t v) {
  while (v >= 0x80) {
    out->push_back((char)((v & 0x7F) | 0x80));
    v >>= 7;
  }
  out->push_back((char)v);
}

void put_tag(std::string* out, int num, int wt) { put_varint(out, ((uint64_t)num << 3) | (uint64_t)wt); }

void put_doubles(std::string* out, int num, const double* v, int n) {
  for (int k = 0; k < n; ++k) {
    put_tag(out, num, 1);
    char b[8];
    std::memcpy(b, &v[k], 8);
    out->append(b, 8);
  }
}

void put_message(std::string* out, int num, const std::string& body) {
  put_tag(out, num, 2);
  put_varint(out, body.size());
  out->append(body);
}

// ------------------------------------------------------------ YAML subset

struct YNode {
  enum Kind { kNull, kScalar, kMap, kSeq } kind = kNull;
  std::string scalar;
  std::vector<std::pair<std::string, std::shared_ptr<YNode>>> map;
  std::vector<std::shared_ptr<YNode>> seq;
  const YNode* get(const std::string& k) const {
    if (kind != kMap) return nullptr;
    for (const auto& kv : map)
      if (kv.first == k) return kv.second.get();
    return nullptr;
  }
};

struct YLine {
  int indent;
  std::string text;   // without indentation and comment
  int open_brackets;  // '[' minus ']' outside quotes
};

int bracket_balance(const std::string& s) {
  bool sq = false, dq = false;
  int b = 0;
  for (char ch : s) {
    if (ch == '\'' && !dq) sq = !sq;
    if (ch == '"' && !sq) dq = !dq;
    if (!sq && !dq) b += (ch == '[') - (ch == ']');
  }
  return b;
}

std::string trim(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && (s[b] == ' ' || s[b] == '\t' || s[b] == '\r')) ++b;
  while (e > b && (s[e - 1] == ' ' || s[e - 1] == '\t' || s[e - 1] == '\r')) --e;
  return s.substr(b, e - b);
}

std::string unquote(const std::string& s) {
  if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
    return s.substr(1, s.size() - 2);
  return s;
}

std::string strip_comment(const std::string& s) {
  bool sq = false, dq = false;
  for (size_t i = 0; i < s.size(); ++i) {
    if (s[i] == '\'' && !dq) sq = !sq;
    if (s[i] == '"' && !sq) dq = !dq;
    if (s[i] == '#' && !sq && !dq && (i == 0 || s[i - 1] == ' ')) return s.substr(0, i);
  }
  return s;
}

std::shared_ptr<YNode> scalar_node(const std::string& text) {
  auto n = std::make_shared<YNode>();
  const std::string t = trim(text);
  if (t.size() >= 2 && t.front() == '[' && t.back() == ']') {
    n->kind = YNode::kSeq;
    std::string inner = t.substr(1, t.size() - 2);
    std::stringstream ss(inner);
    for (std::string item; std::getline(ss, item, ',');) {
      const std::string v = trim(item);
      if (v.empty()) continue;
      auto c = std::make_shared<YNode>();
      c->kind = YNode::kScalar;
      c->scalar = unquote(v);
      n->seq.push_back(c);
    }
    return n;
  }
  if (t == "{}") { n->kind = YNode::kMap; return n; }
  if (t.empty() || t == "~" || t == "null") return n;
  n->kind = YNode::kScalar;
  n->scalar = unquote(t);
  return n;
}

// Position of the "key:" separator of a mapping line, or npos.
size_t key_colon(const std::string& t) {
  bool sq = false, dq = false;
  for (size_t i = 0; i < t.size(); ++i) {
    if (t[i] == '\'' && !dq) sq = !sq;
    if (t[i] == '"' && !sq) dq = !dq;
    if (t[i] == '[' && !sq && !dq) return std::string::npos;
    if (t[i] == ':' && !sq && !dq && (i + 1 == t.size() || t[i + 1] == ' ')) return i;
  }
  return std::string::npos;
}

std::shared_ptr<YNode> parse_block(const std::vector<YLine>& L, size_t& i, int indent);

// Value after "key:" (rest of the line, or a nested block on the next lines).
std::shared_ptr<YNode> parse_value(const std::vector<YLine>& L, size_t& i, const std::string& rest, int indent) {
  if (!trim(rest).empty()) return scalar_node(rest);
  if (i < L.size() && (L[i].indent > indent || (L[i].indent == indent && L[i].text.rfind("- ", 0) == 0) ||
                       (L[i].indent == indent && L[i].text == "-")))
    return parse_block(L, i, L[i].indent);
  return std::make_shared<YNode>();
}

std::shared_ptr<YNode> parse_block(const std::vector<YLine>& L, size_t& i, int indent) {
  auto n = std::make_shared<YNode>();
  if (i >= L.size()) return n;
  const bool seq = L[i].text == "-" || L[i].text.rfind("- ", 0) == 0;
  n->kind = seq ? YNode::kSeq : YNode::kMap;
  while (i < L.size() && L[i].indent == indent) {
    const std::string t = L[i].text;
    if (seq) {
      if (!(t == "-" || t.rfind("- ", 0) == 0)) break;
      const std::string rest = t.size() > 1 ? t.substr(2) : std::string();
      ++i;
      const std::string r = trim(rest);
      if (r.empty()) {
        n->seq.push_back(i < L.size() && L[i].indent > indent ? parse_block(L, i, L[i].indent)
                                                              : std::make_shared<YNode>());
      } else if (key_colon(r) != std::string::npos) {
        // "- key: value" starts a mapping whose further keys are indented to the key.
        auto m = std::make_shared<YNode>();
        m->kind = YNode::kMap;
        const size_t c = key_colon(r);
        const std::string key = unquote(trim(r.substr(0, c)));
        m->map.push_back({key, parse_value(L, i, r.substr(c + 1), indent + 2)});
        while (i < L.size() && L[i].indent == indent + 2 && L[i].text.rfind("- ", 0) != 0) {
          const std::string tt = L[i].text;
          const size_t cc = key_colon(tt);
          if (cc == std::string::npos) break;
          const std::string k2 = unquote(trim(tt.substr(0, cc)));
          ++i;
          m->map.push_back({k2, parse_value(L, i, tt.substr(cc + 1), indent + 2)});
        }
        n->seq.push_back(m);
      } else {
        n->seq.push_back(scalar_node(r));
      }
    } else {
      const size_t c = key_colon(t);
      if (c == std::string::npos) break;
      const std::string key = unquote(trim(t.substr(0, c)));
      ++i;
      n->map.push_back({key, parse_value(L, i, t.substr(c + 1), indent)});
    }
  }
  return n;
}

std::shared_ptr<YNode> parse_yaml(const std::string& text) {
  std::vector<YLine> lines;
  std::stringstream ss(text);
  for (std::string raw; std::getline(ss, raw);) {
    const std::string s = strip_comment(raw);
    if (trim(s).empty() || trim(s) == "---") continue;
    int ind = 0;
    while (ind < (int)s.size() && s[ind] == ' ') ++ind;
    // A flow sequence may wrap onto continuation lines: join them.
    if (!lines.empty() && lines.back().open_brackets > 0) {
      lines.back().text += " " + trim(s);
      lines.back().open_brackets += bracket_balance(s);
      continue;
    }
    lines.push_back(YLine{ind, trim(s), bracket_balance(s)});
  }
  size_t i = 0;
  return lines.empty() ? std::make_shared<YNode>() : parse_block(lines, i, lines[0].indent);
}

bool to_double(const YNode* n, double* v) {
  if (!n || n->kind != YNode::kScalar) return false;
  char* end = nullptr;
  *v = std::strtod(n->scalar.c_str(), &end);
  return end && *end == '\0' && end != n->scalar.c_str();
}

// A YAML matrix {rows, cols, data} or a plain list, row-major.
bool yaml_matrix(const YNode* n, std::vector<double>* out, int* rows, int* cols) {
  out->clear();
  const YNode* data = n;
  double r = -1, c = -1;
  if (n && n->kind == YNode::kMap) {
    data = n->get("data");
    if (!to_double(n->get("rows"), &r) || !to_double(n->get("cols"), &c)) return false;
  }
  if (!data || data->kind != YNode::kSeq) return false;
  for (const auto& e : data->seq) {
    double v;
    if (!to_double(e.get(), &v)) return false;
    out->push_back(v);
  }
  *rows = r >= 0 ? (int)r : (int)out->size();
  *cols = c >= 0 ? (int)c : 1;
  return (size_t)(*rows) * (size_t)(*cols) == out->size();
}

Key hex_key(const std::string& h) {
  if (h.size() != 32) return Key(0, 0);
  uint8_t b[16];
  for (int k = 0; k < 16; ++k) b[k] = (uint8_t)std::strtoul(h.substr(2 * k, 2).c_str(), nullptr, 16);
  uint64_t a = 0, c = 0;
  std::memcpy(&a, b, 8);
  std::memcpy(&c, b + 8, 8);
  return Key(a, c);  // the 16 bytes in memory order, little-endian words
}

// ------------------------------------------------------------ rotations

// Rotation matrix -> Hamilton quaternion (x, y, z, w), w >= 0.
void matrix_to_quat(const double R[9], double q[4]) {
  const double t = R[0] + R[4] + R[8];
  if (t > 0) {
    const double s = 2.0 * std::sqrt(t + 1.0);
    q[0] = (R[7] - R[5]) / s;
    q[1] = (R[2] - R[6]) / s;
    q[2] = (R[3] - R[1]) / s;
    q[3] = 0.25 * s;
  } else {
    int i = 0;
    if (R[4] > R[0]) i = 1;
    if (R[8] > R[3 * i + i]) i = 2;
    const int j = (i + 1) % 3, k = (i + 2) % 3;
    const double s = 2.0 * std::sqrt(1.0 + R[3 * i + i] - R[3 * j + j] - R[3 * k + k]);
    q[i] = 0.25 * s;
    q[j] = (R[3 * j + i] + R[3 * i + j]) / s;
    q[k] = (R[3 * k + i] + R[3 * i + k]) / s;
    q[3] = (R[3 * k + j] - R[3 * j + k]) / s;
  }
  const double nrm = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  const double sg = q[3] < 0 ? -1.0 : 1.0;
  for (int a = 0; a < 4; ++a) q[a] = sg * q[a] / nrm;
}

// ------------------------------------------------------------ the map

struct ProtoFile {
  int32_t file_type = 0;
  std::string path;
  bool zipped = false;
  std::string bytes;  // decompressed VIMap message
};

struct Camera {
  double intr[4], dist[4];
  int32_t dist_type, width, height, model;
  double xi;
  double T_B_C[16];
};

}  // namespace

struct mlba_vimap {
  std::string folder;
  std::vector<ProtoFile> files;
  std::vector<Key> vertex_ids, landmark_ids;
  std::map<Key, int32_t> vindex, lindex_store;  // vertex id -> index; landmark id -> index (last store wins)
  // Problem arrays.
  std::vector<double> vertices, baseframes, landmarks, cam_q, intr, dist, kp, sigma, xi;
  std::vector<int32_t> dist_type, resolution, vertex_mission, store, ov, ol, oc, model;
  std::vector<uint8_t> vconst, lconst, quality;
  bool any_upm = false;
};

namespace {

int load_cameras(const std::string& path, std::map<Key, std::vector<Camera>>* out) {
  std::string text;
  if (!read_file(path, &text)) return fail(MLBA_EINVAL, "cannot read " + path);
  auto root = parse_yaml(text);
  const YNode* sensors = root->get("sensors");
  if (!sensors || sensors->kind != YNode::kSeq) return MLBA_OK;
  static const std::pair<const char*, int> kDist[] = {
      {"none", MLBA_DISTORTION_NONE}, {"radial-tangential", MLBA_DISTORTION_RADTAN}, {"equidistant", MLBA_DISTORTION_EQUIDISTANT},
      {"fisheye", MLBA_DISTORTION_FISHEYE}};
  for (const auto& sp : sensors->seq) {
    const YNode* s = sp.get();
    const YNode* st = s->get("sensor_type");
    if (!st || st->kind != YNode::kScalar) continue;
    std::string stype = st->scalar;
    std::transform(stype.begin(), stype.end(), stype.begin(), ::toupper);
    if (stype != "NCAMERA") continue;
    const YNode* id = s->get("id");
    const YNode* cams = s->get("cameras");
    std::vector<Camera> list;
    if (cams && cams->kind == YNode::kSeq) {
      for (const auto& cp : cams->seq) {
        const YNode* cam = cp->get("camera");
        if (!cam) return fail(MLBA_EINVAL, "sensors.yaml: camera entry without 'camera'");
        Camera c{};
        const YNode* ty = cam->get("type");
        const std::string ctype = ty && ty->kind == YNode::kScalar ? ty->scalar : "pinhole";
        if (ctype != "pinhole" && ctype != "unified-projection")
          return fail(MLBA_EINVAL, "camera type '" + ctype + "' is not supported (pinhole, unified-projection)");
        std::vector<double> iv;
        int r, cc;
        if (!yaml_matrix(cam->get("intrinsics"), &iv, &r, &cc)) return fail(MLBA_EINVAL, "camera intrinsics");
        c.model = MLBA_CAMERA_PINHOLE;
        c.xi = 0.0;
        size_t off = 0;
        if (ctype == "unified-projection") {
          if (iv.empty()) return fail(MLBA_EINVAL, "camera intrinsics");
          c.model = MLBA_CAMERA_UNIFIED_PROJECTION;
          c.xi = iv[0];
          off = 1;
        }
        if (iv.size() < off + 4) return fail(MLBA_EINVAL, "camera intrinsics");
        for (int k = 0; k < 4; ++k) c.intr[k] = iv[off + k];
        std::string dname = "none";
        const YNode* d = cam->get("distortion");
        if (d && d->get("type") && d->get("type")->kind == YNode::kScalar) dname = d->get("type")->scalar;
        c.dist_type = -1;
        for (const auto& kv : kDist)
          if (dname == kv.first) c.dist_type = kv.second;
        if (c.dist_type < 0) return fail(MLBA_EINVAL, "distortion '" + dname + "' is not supported");
        for (int k = 0; k < 4; ++k) c.dist[k] = 0.0;
        if (dname != "none") {
          std::vector<double> dv;
          if (!yaml_matrix(d->get("parameters"), &dv, &r, &cc) || dv.size() > 4)
            return fail(MLBA_EINVAL, "distortion parameters");
          for (size_t k = 0; k < dv.size(); ++k) c.dist[k] = dv[k];
        }
        double w, h;
        if (!to_double(cam->get("image_width"), &w) || !to_double(cam->get("image_height"), &h))
          return fail(MLBA_EINVAL, "camera image size");
        c.width = (int32_t)w;
        c.height = (int32_t)h;
        std::vector<double> T;
        if (cp->get("T_B_C")) {
          if (!yaml_matrix(cp->get("T_B_C"), &T, &r, &cc) || r != 4 || cc != 4) return fail(MLBA_EINVAL, "T_B_C");
          for (int k = 0; k < 16; ++k) c.T_B_C[k] = T[k];
        } else if (cp->get("T_C_B")) {
          if (!yaml_matrix(cp->get("T_C_B"), &T, &r, &cc) || r != 4 || cc != 4) return fail(MLBA_EINVAL, "T_C_B");
          // Inverse of the rigid transform [R p; 0 1]: [R' -R'p; 0 1].
          for (int a = 0; a < 3; ++a) {
            for (int b = 0; b < 3; ++b) c.T_B_C[4 * a + b] = T[4 * b + a];
            c.T_B_C[4 * a + 3] = -(T[a] * T[3] + T[4 + a] * T[7] + T[8 + a] * T[11]);
          }
          c.T_B_C[12] = c.T_B_C[13] = c.T_B_C[14] = 0.0;
          c.T_B_C[15] = 1.0;
        } else {
          return fail(MLBA_EINVAL, "camera without T_B_C / T_C_B");
        }
        list.push_back(c);
      }
    }
    (*out)[hex_key(id && id->kind == YNode::kScalar ? id->scalar : std::string())] = list;
  }
  return MLBA_OK;
}

struct VertexInfo {
  Span span;
  Key mission;
  std::vector<Key> outgoing;
};

int load(const std::string& folder, mlba_vimap* m) {
  const std::string base = folder + "/vi_map/";
  auto read_proto = [&](const std::string& path, std::string* out, bool* zipped) -> bool {
    std::string data;
    if (!read_file(path, &data)) return false;
    *zipped = data.size() >= 2 && (uint8_t)data[0] == 0x1f && (uint8_t)data[1] == 0x8b;
    if (*zipped) return gunzip(data, out);
    *out = std::move(data);
    return true;
  };
  std::string meta;
  bool zm;
  if (!read_proto(base + "vi_map_metadata", &meta, &zm)) return fail(MLBA_EINVAL, "cannot read " + base + "vi_map_metadata");
  {
    Reader r(Span{reinterpret_cast<const uint8_t*>(meta.data()), meta.size()});
    Field f;
    while (r.next(&f)) {
      if (f.num != 1 || f.wt != 2) continue;
      ProtoFile pf;
      Reader rr(f.s);
      Field g;
      while (rr.next(&g)) {
        if (g.num == 1 && g.wt == 0) pf.file_type = (int32_t)g.v;
        if (g.num == 2 && g.wt == 2) pf.path.assign(reinterpret_cast<const char*>(g.s.p), g.s.n);
      }
      m->files.push_back(std::move(pf));
    }
    if (!r.ok()) return fail(MLBA_EINVAL, "malformed vi_map_metadata");
  }
  for (ProtoFile& pf : m->files)
    if (!read_proto(base + pf.path, &pf.bytes, &pf.zipped)) return fail(MLBA_EINVAL, "cannot read " + base + pf.path);
  // Missions, baseframes, vertices, edges (first-seen order, later values win).
  struct Mission { Key baseframe, root, ncamera; };
  std::vector<Key> mission_order;
  std::map<Key, Mission> missions;
  std::map<Key, std::vector<double>> baseframes;
  std::vector<Key> vertex_order;
  std::map<Key, VertexInfo> vertices;
  std::map<Key, Key> edge_to;  // edge id -> viwls 'to' vertex
  for (const ProtoFile& pf : m->files) {
    const Span all{reinterpret_cast<const uint8_t*>(pf.bytes.data()), pf.bytes.size()};
    std::vector<Key> ids[5];  // vertex, edge, mission, baseframe (fields 1, 3, 5, 7)
    std::vector<Span> items[5];
    Reader r(all);
    Field f;
    while (r.next(&f)) {
      if (f.wt != 2 || f.num < 1 || f.num > 8) continue;
      const int slot = (f.num - 1) / 2;
      if ((f.num & 1)) ids[slot].push_back(parse_id(f.s));
      else items[slot].push_back(f.s);
    }
    if (!r.ok()) return fail(MLBA_EINVAL, "malformed VIMap file " + pf.path);
    for (size_t k = 0; k < items[2].size() && k < ids[2].size(); ++k) {
      Mission ms{};
      Reader rr(items[2][k]);
      Field g;
      while (rr.next(&g)) {
        if (g.wt != 2) continue;
        if (g.num == 1) ms.baseframe = parse_id(g.s);
        if (g.num == 2) ms.root = parse_id(g.s);
        if (g.num == 8) ms.ncamera = parse_id(g.s);
      }
      if (!missions.count(ids[2][k])) mission_order.push_back(ids[2][k]);
      missions[ids[2][k]] = ms;
    }
    for (size_t k = 0; k < items[3].size() && k < ids[3].size(); ++k) {
      std::vector<double> t;
      Reader rr(items[3][k]);
      Field g;
      while (rr.next(&g))
        if (g.num == 1) add_doubles(g, &t);
      baseframes[ids[3][k]] = t;
    }
    for (size_t k = 0; k < items[0].size() && k < ids[0].size(); ++k) {
      VertexInfo vi;
      vi.span = items[0][k];
      Reader rr(items[0][k]);
      Field g;
      while (rr.next(&g)) {
        if (g.num == 2 && g.wt == 2) vi.outgoing.push_back(parse_id(g.s));
        if (g.num == 14 && g.wt == 2) vi.mission = parse_id(g.s);
      }
      if (!vertices.count(ids[0][k])) vertex_order.push_back(ids[0][k]);
      vertices[ids[0][k]] = vi;
    }
    for (size_t k = 0; k < items[1].size() && k < ids[1].size(); ++k) {
      Reader rr(items[1][k]);
      Field g;
      while (rr.next(&g)) {
        if (g.num != 1 || g.wt != 2) continue;  // Edge.viwls
        Reader re(g.s);
        Field h;
        while (re.next(&h))
          if (h.num == 2 && h.wt == 2) edge_to[ids[1][k]] = parse_id(h.s);
      }
    }
  }
  std::map<Key, std::vector<Camera>> ncams;
  int rc = load_cameras(base + "sensors.yaml", &ncams);
  if (rc) return rc;
  // Vertices mission by mission along the VIWLS chain from the root, then the
  // mission's other vertices in file order.
  std::vector<Key> order;
  std::vector<int32_t> mission_of;
  for (size_t mi = 0; mi < mission_order.size(); ++mi) {
    const Key mk = mission_order[mi];
    std::map<Key, bool> seen;
    std::vector<Key> chain;
    Key cur = missions[mk].root;
    while (vertices.count(cur) && !seen.count(cur)) {
      seen[cur] = true;
      chain.push_back(cur);
      Key nxt(0, 0);
      bool found = false;
      for (const Key& eid : vertices[cur].outgoing) {
        auto it = edge_to.find(eid);
        if (it != edge_to.end()) { nxt = it->second; found = true; break; }
      }
      if (!found) break;
      cur = nxt;
    }
    for (const Key& k : chain) { order.push_back(k); mission_of.push_back((int32_t)mi); }
    for (const Key& k : vertex_order)
      if (vertices[k].mission == mk && !seen.count(k)) { order.push_back(k); mission_of.push_back((int32_t)mi); }
  }
  m->vertex_ids = order;
  for (size_t i = 0; i < order.size(); ++i) m->vindex[order[i]] = (int32_t)i;
  // Cameras: the NCamera of every mission, in mission order.
  std::map<Key, int32_t> cam_base;
  std::vector<Camera> cams;
  for (const Key& mk : mission_order) {
    const Key nk = missions[mk].ncamera;
    if (cam_base.count(nk)) continue;
    auto it = ncams.find(nk);
    if (it == ncams.end()) return fail(MLBA_EINVAL, "a mission has no NCAMERA sensor in sensors.yaml");
    cam_base[nk] = (int32_t)cams.size();
    cams.insert(cams.end(), it->second.begin(), it->second.end());
  }
  // Landmarks from the stores, then observations from the frames.
  const size_t P = order.size();
  for (size_t vi = 0; vi < P; ++vi) {
    Reader r(vertices[order[vi]].span);
    Field f;
    while (r.next(&f)) {
      if (f.num != 8 || f.wt != 2) continue;
      Reader rs(f.s);
      Field g;
      while (rs.next(&g)) {
        if (g.num != 1 || g.wt != 2) continue;
        Key id(0, 0);
        std::vector<double> pos;
        int32_t q = 0;
        Reader rl(g.s);
        Field h;
        while (rl.next(&h)) {
          if (h.num == 1 && h.wt == 2) id = parse_id(h.s);
          if (h.num == 2) add_doubles(h, &pos);
          if (h.num == 7 && h.wt == 0) q = (int32_t)h.v;
        }
        if (pos.size() != 3) return fail(MLBA_EINVAL, "landmark position is not 3 doubles");
        m->lindex_store[id] = (int32_t)m->landmark_ids.size();
        m->landmark_ids.push_back(id);
        m->landmarks.insert(m->landmarks.end(), pos.begin(), pos.end());
        m->store.push_back((int32_t)vi);
        m->quality.push_back((uint8_t)q);
      }
    }
  }
  for (size_t vi = 0; vi < P; ++vi) {
    const int32_t cb = cam_base[missions[mission_order[mission_of[vi]]].ncamera];
    std::vector<double> T;
    Reader r(vertices[order[vi]].span);
    Field f;
    int fi_count = 0;
    while (r.next(&f)) {
      if (f.num == 3) add_doubles(f, &T);
      if (f.num != 7 || f.wt != 2) continue;
      Reader rn(f.s);
      Field g;
      while (rn.next(&g)) {
        if (g.num != 2 || g.wt != 2) continue;  // VisualNFrame.frames
        const int fi = fi_count++;
        std::vector<double> meas, sig;
        std::vector<Key> lids;
        Reader rf(g.s);
        Field h;
        while (rf.next(&h)) {
          if (h.num == 3) add_doubles(h, &meas);
          if (h.num == 4) add_doubles(h, &sig);
          if (h.num == 7 && h.wt == 2) lids.push_back(parse_id(h.s));
        }
        for (size_t k = 0; k < lids.size(); ++k) {
          auto it = m->lindex_store.find(lids[k]);
          if (it == m->lindex_store.end()) continue;
          if (2 * k + 1 >= meas.size() || k >= sig.size()) return fail(MLBA_EINVAL, "frame keypoints and landmark ids differ");
          m->ov.push_back((int32_t)vi);
          m->ol.push_back(it->second);
          m->oc.push_back(cb + fi);
          m->kp.push_back(meas[2 * k]);
          m->kp.push_back(meas[2 * k + 1]);
          m->sigma.push_back(sig[k]);
        }
      }
    }
    if (T.size() != 7) return fail(MLBA_EINVAL, "vertex T_M_I is not 7 doubles");
    const double s = T[3] < 0 ? -1.0 : 1.0;
    for (int a = 0; a < 4; ++a) m->vertices.push_back(s * T[a]);
    for (int a = 4; a < 7; ++a) m->vertices.push_back(T[a]);
  }
  for (const Key& mk : mission_order) {
    auto it = baseframes.find(missions[mk].baseframe);
    if (it == baseframes.end() || it->second.size() != 7) return fail(MLBA_EINVAL, "mission baseframe T_G_M missing");
    const std::vector<double>& t = it->second;
    double q[4] = {-t[0], -t[1], -t[2], t[3]};  // JPL q_GM = inverse of the Hamilton q_G_M
    const double s = q[3] < 0 ? -1.0 : 1.0;
    for (int a = 0; a < 4; ++a) m->baseframes.push_back(s * q[a]);
    for (int a = 4; a < 7; ++a) m->baseframes.push_back(t[a]);
  }
  for (const Camera& c : cams) {
    double R[9], q[4];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) R[3 * a + b] = c.T_B_C[4 * a + b];
    matrix_to_quat(R, q);  // JPL q_CI = inverse of the Hamilton q_C_I = q(R_B_C)
    for (int a = 0; a < 4; ++a) m->cam_q.push_back(q[a]);
    for (int a = 0; a < 3; ++a)  // C_p_CI = -R_B_C' p_B_C
      m->cam_q.push_back(-(R[a] * c.T_B_C[3] + R[3 + a] * c.T_B_C[7] + R[6 + a] * c.T_B_C[11]));
    m->intr.insert(m->intr.end(), c.intr, c.intr + 4);
    m->dist.insert(m->dist.end(), c.dist, c.dist + 4);
    m->dist_type.push_back(c.dist_type);
    m->resolution.push_back(c.width);
    m->resolution.push_back(c.height);
    m->model.push_back(c.model);
    m->xi.push_back(c.xi);
    m->any_upm = m->any_upm || c.model != MLBA_CAMERA_PINHOLE;
  }
  m->vertex_mission = mission_of;
  m->vconst.assign(P, 0);
  if (P) m->vconst[0] = 1;
  m->lconst.assign(m->landmark_ids.size(), 0);
  return MLBA_OK;
}

// ------------------------------------------------------------ saving

struct SaveState {
  const mlba_vimap* m;
  const double* vertices;
  const double* landmarks;
  const uint8_t* quality;
};

std::string rewrite_landmark(Span s, const SaveState& st) {
  Key id(0, 0);
  {
    Reader r(s);
    Field f;
    while (r.next(&f))
      if (f.num == 1 && f.wt == 2) id = parse_id(f.s);
  }
  auto it = st.m->lindex_store.find(id);
  const int l = it == st.m->lindex_store.end() ? -1 : it->second;
  std::string out;
  bool pos_done = false, q_done = false;
  const bool set_q = l >= 0 && st.quality && (st.quality[l] == 1 || st.quality[l] == 2);
  Reader r(s);
  Field f;
  while (r.next(&f)) {
    if (l >= 0 && f.num == 2) {
      if (!pos_done) put_doubles(&out, 2, st.landmarks + 3 * l, 3);
      pos_done = true;
    } else if (set_q && f.num == 7) {
      if (!q_done) { put_tag(&out, 7, 0); put_varint(&out, st.quality[l]); }
      q_done = true;
    } else {
      out.append(reinterpret_cast<const char*>(f.raw.p), f.raw.n);
    }
  }
  if (l >= 0 && !pos_done) put_doubles(&out, 2, st.landmarks + 3 * l, 3);
  if (set_q && !q_done) { put_tag(&out, 7, 0); put_varint(&out, st.quality[l]); }
  return out;
}

std::string rewrite_vertex(Span s, int v, const SaveState& st) {
  std::string out;
  bool t_done = false;
  Reader r(s);
  Field f;
  while (r.next(&f)) {
    if (v >= 0 && f.num == 3) {
      if (!t_done) put_doubles(&out, 3, st.vertices + 7 * v, 7);
      t_done = true;
    } else if (f.num == 8 && f.wt == 2) {  // landmark_store
      std::string store;
      Reader rs(f.s);
      Field g;
      while (rs.next(&g)) {
        if (g.num == 1 && g.wt == 2) put_message(&store, 1, rewrite_landmark(g.s, st));
        else store.append(reinterpret_cast<const char*>(g.raw.p), g.raw.n);
      }
      put_message(&out, 8, store);
    } else {
      out.append(reinterpret_cast<const char*>(f.raw.p), f.raw.n);
    }
  }
  if (v >= 0 && !t_done) put_doubles(&out, 3, st.vertices + 7 * v, 7);
  return out;
}

std::string rewrite_file(const ProtoFile& pf, const SaveState& st) {
  const Span all{reinterpret_cast<const uint8_t*>(pf.bytes.data()), pf.bytes.size()};
  std::vector<Key> vids;
  {
    Reader r(all);
    Field f;
    while (r.next(&f))
      if (f.num == 1 && f.wt == 2) vids.push_back(parse_id(f.s));
  }
  std::string out;
  size_t k = 0;
  Reader r(all);
  Field f;
  while (r.next(&f)) {
    if (f.num == 2 && f.wt == 2) {
      int v = -1;
      if (k < vids.size()) {
        auto it = st.m->vindex.find(vids[k]);
        if (it != st.m->vindex.end()) v = it->second;
      }
      ++k;
      put_message(&out, 2, rewrite_vertex(f.s, v, st));
    } else {
      out.append(reinterpret_cast<const char*>(f.raw.p), f.raw.n);
    }
  }
  return out;
}

bool copy_file(const std::string& from, const std::string& to) {
  if (from == to) return true;
  std::string data;
  return read_file(from, &data) && write_file(to, data);
}

}  // namespace

extern "C" {

int32_t mlba_vimap_load(const char* folder, mlba_vimap** out) {
  if (!folder || !out) return fail(MLBA_EINVAL, "NULL argument");
  *out = nullptr;
  auto m = std::make_unique<mlba_vimap>();
  m->folder = folder;
  const int rc = load(folder, m.get());
  if (rc) return rc;
  *out = m.release();
  return MLBA_OK;
}

void mlba_vimap_free(mlba_vimap* m) { delete m; }

int32_t mlba_vimap_describe(const mlba_vimap* m, mlba_problem_desc* d) {
  if (!m || !d) return fail(MLBA_EINVAL, "NULL argument");
  std::memset(d, 0, sizeof(*d));
  d->num_vertices = (int32_t)m->vertex_ids.size();
  d->num_landmarks = (int32_t)m->landmark_ids.size();
  d->num_observations = (int32_t)m->ov.size();
  d->num_cameras = (int32_t)m->dist_type.size();
  d->num_missions = (int32_t)(m->baseframes.size() / 7);
  d->vertex_q_IM__M_p_MI = m->vertices.data();
  d->baseframe_q_GM__G_p_GM = m->baseframes.data();
  d->landmark_p_B = m->landmarks.data();
  d->camera_q_CI__C_p_CI = m->cam_q.data();
  d->camera_intrinsics = m->intr.data();
  d->camera_distortion = m->dist.data();
  d->camera_distortion_type = m->dist_type.data();
  d->camera_resolution = m->resolution.data();
  d->vertex_mission = m->vertex_mission.data();
  d->landmark_store_vertex = m->store.data();
  d->obs_vertex = m->ov.data();
  d->obs_landmark = m->ol.data();
  d->obs_camera = m->oc.data();
  d->obs_keypoint = m->kp.data();
  d->obs_sigma = m->sigma.data();
  d->vertex_constant = m->vconst.data();
  d->landmark_constant = m->lconst.data();
  d->camera_model = m->any_upm ? m->model.data() : nullptr;
  d->camera_xi = m->any_upm ? m->xi.data() : nullptr;
  return MLBA_OK;
}

int32_t mlba_vimap_landmark_quality(const mlba_vimap* m, uint8_t* quality) {
  if (!m || !quality) return fail(MLBA_EINVAL, "NULL argument");
  std::copy(m->quality.begin(), m->quality.end(), quality);
  return MLBA_OK;
}

int32_t mlba_vimap_save(const mlba_vimap* m, const char* folder, const double* vertices, const double* landmarks,
                        const uint8_t* landmark_quality) {
  if (!m || !folder || !vertices || !landmarks) return fail(MLBA_EINVAL, "NULL argument");
  const std::string src = m->folder + "/vi_map/", dst = std::string(folder) + "/vi_map/";
  {
    // The destination folder must exist (mkdir -p semantics are left to the caller).
    std::ofstream probe(dst + ".mlba_probe");
    if (!probe) return fail(MLBA_EINVAL, "cannot write into " + dst);
  }
  std::remove((dst + ".mlba_probe").c_str());
  SaveState st{m, vertices, landmarks, landmark_quality};
  for (const ProtoFile& pf : m->files) {
    std::string data = rewrite_file(pf, st);
    if (pf.zipped) {
      std::string z;
      if (!gzip(data, &z)) return fail(MLBA_EDEVICE, "gzip failed");
      data.swap(z);
    }
    if (!write_file(dst + pf.path, data)) return fail(MLBA_EINVAL, "cannot write " + dst + pf.path);
  }
  if (!copy_file(src + "vi_map_metadata", dst + "vi_map_metadata") ||
      !copy_file(src + "sensors.yaml", dst + "sensors.yaml"))
    return fail(MLBA_EINVAL, "cannot copy the metadata or sensors.yaml into " + dst);
  return MLBA_OK;
}

const char* mlba_vimap_last_error(void) { return g_error.c_str(); }

}  // extern "C"
