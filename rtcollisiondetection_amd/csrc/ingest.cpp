// Geometry ingestion on the host: OBJ meshes, URDF kinematic chains and sphere sets, so that a robot
// or an obstacle is loaded from the same files RTCD reads.
//
// * OBJ (RTCD/Meshes/mesh.h:311-365, meshModel::loadOBJ through tinyobjloader, fetched unpinned from
//   its release branch by cmake/configure_tiny_obj_loader.cmake): `v` and `f` records; quads split
//   along the shorter diagonal as that loader does, larger polygons fan-triangulated (the loader's
//   newer releases use earcut there, so those faces are not pinned to it); vertices deduplicated by
//   exact value in first-use order (mesh.h:218-231, addVertex with float3less: -0 and +0 are one
//   vertex).
// * edges (mesh.h:200-216 addEdge, 378-620): unique undirected edges in first-seen order, or the loop
//   edges (three per triangle).  Keys are 64-bit (the reference's int Cantor hash overflows above
//   ~32k vertices).
// * URDF (scripts/URDFParser/urdfParserUtils.py:527-605 generate_masks_for_joint, whose output is
//   RTCD/Robot/models/panda.h:16-44): the serial chain from the root link; each revolute or continuous
//   joint with origin (xyz, rpy) and unit axis k gives T_local(q) = cos q C + sin q S + O with
//   C = R0 (I - k k^T), S = R0 [k]x, O = [R0 k k^T | xyz] (Rodrigues), entries below 1e-9 set to 0.
//   Fixed joints before a movable joint fold into its origin.  Link meshes are the first <visual>
//   mesh of each link, moved by the visual origin and scaled.
// * sphere sets (mesh.h:825-848): text lines `x y z r` in the link frame.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/rtcd.h"

namespace rtcd {
extern thread_local std::string g_err;  // the library's last-error string (rtcd_last_error)
}

namespace {

constexpr double kPi = 3.14159265358979323846;

rtcd_status set_err(rtcd_status st, const std::string& msg) {
    rtcd::g_err = msg;
    return st;
}

std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::invalid_argument("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// ------------------------------------------------------------------------------------------ OBJ
struct Mesh {
    std::vector<float> v;    // [V][3]
    std::vector<int32_t> f;  // [F][3]
};

struct VKey {
    float x, y, z;
    bool operator<(const VKey& o) const {
        if (x != o.x) return x < o.x;
        if (y != o.y) return y < o.y;
        return z < o.z;
    }
};

Mesh parse_obj(const std::string& text, const std::string& name) {
    std::vector<float> raw;
    Mesh m;
    std::map<VKey, int32_t> known;
    std::vector<int32_t> remap;  // raw vertex -> deduplicated vertex (-1: not used yet)
    auto use = [&](int32_t i) {
        if (remap[(size_t) i] < 0) {
            const VKey k{raw[(size_t) i * 3], raw[(size_t) i * 3 + 1], raw[(size_t) i * 3 + 2]};
            auto it = known.find(k);
            if (it == known.end()) {
                it = known.emplace(k, (int32_t) (m.v.size() / 3)).first;
                m.v.insert(m.v.end(), {k.x, k.y, k.z});
            }
            remap[(size_t) i] = it->second;
        }
        return remap[(size_t) i];
    };
    std::istringstream in(text);
    std::string line;
    size_t lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        if (line.size() < 2) continue;
        if (line[0] == 'v' && (line[1] == ' ' || line[1] == '\t')) {
            std::istringstream ls(line.substr(2));
            float x, y, z;
            if (!(ls >> x >> y >> z)) throw std::invalid_argument(name + ":" + std::to_string(lineno) + ": bad vertex");
            raw.insert(raw.end(), {x, y, z});
            remap.push_back(-1);
        } else if (line[0] == 'f' && (line[1] == ' ' || line[1] == '\t')) {
            std::istringstream ls(line.substr(2));
            std::string tok;
            std::vector<int32_t> idx;
            while (ls >> tok) {
                const long i = std::strtol(tok.c_str(), nullptr, 10);  // "v", "v/vt", "v//vn", "v/vt/vn"
                const long nv = (long) remap.size();
                const long r = i > 0 ? i - 1 : nv + i;
                if (i == 0 || r < 0 || r >= nv)
                    throw std::invalid_argument(name + ":" + std::to_string(lineno) + ": face index out of range");
                idx.push_back((int32_t) r);
            }
            if (idx.size() < 3) throw std::invalid_argument(name + ":" + std::to_string(lineno) + ": face with < 3 vertices");
            if (idx.size() == 4) {
                // a quad splits along its shorter diagonal, (0,1,2)+(0,2,3) or (0,1,3)+(1,2,3), with the
                // squared lengths in float, as tinyobjloader's release branch triangulates quads
                auto sq = [&](int32_t p, int32_t q) {
                    const float dx = raw[(size_t) q * 3] - raw[(size_t) p * 3], dy = raw[(size_t) q * 3 + 1] - raw[(size_t) p * 3 + 1],
                                dz = raw[(size_t) q * 3 + 2] - raw[(size_t) p * 3 + 2];
                    return dx * dx + dy * dy + dz * dz;
                };
                const int32_t q0 = use(idx[0]), q1 = use(idx[1]), q2 = use(idx[2]), q3 = use(idx[3]);
                if (sq(idx[0], idx[2]) < sq(idx[1], idx[3])) m.f.insert(m.f.end(), {q0, q1, q2, q0, q2, q3});
                else m.f.insert(m.f.end(), {q0, q1, q3, q1, q2, q3});
            } else {
                for (size_t k = 1; k + 1 < idx.size(); ++k) {  // fan
                    const int32_t a = use(idx[0]), b = use(idx[k]), c = use(idx[k + 1]);
                    m.f.insert(m.f.end(), {a, b, c});
                }
            }
        }
    }
    return m;
}

// ------------------------------------------------------------------------------------------ XML
// The subset of XML a URDF uses: elements with attributes, nesting, comments, declarations.
struct Node {
    std::string tag;
    std::map<std::string, std::string> attr;
    std::vector<std::unique_ptr<Node>> kids;
    const Node* child(const std::string& t) const {
        for (auto& k : kids)
            if (k->tag == t) return k.get();
        return nullptr;
    }
    std::string get(const std::string& a, const std::string& def = "") const {
        auto it = attr.find(a);
        return it == attr.end() ? def : it->second;
    }
};

std::unique_ptr<Node> parse_xml(const std::string& s) {
    size_t i = 0;
    auto fail = [&](const std::string& what) { throw std::invalid_argument("URDF: " + what + " at byte " + std::to_string(i)); };
    auto ws = [&]() {
        while (i < s.size() && std::isspace((unsigned char) s[i])) ++i;
    };
    auto name = [&]() {
        const size_t b = i;
        while (i < s.size() && (std::isalnum((unsigned char) s[i]) || s[i] == '_' || s[i] == '-' || s[i] == ':' || s[i] == '.')) ++i;
        if (i == b) fail("expected a name");
        return s.substr(b, i - b);
    };
    auto root = std::make_unique<Node>();
    std::vector<Node*> stack{root.get()};
    while (i < s.size()) {
        const size_t lt = s.find('<', i);
        if (lt == std::string::npos) break;
        i = lt + 1;
        if (s.compare(i, 3, "!--") == 0) {
            const size_t e = s.find("-->", i);
            if (e == std::string::npos) fail("unterminated comment");
            i = e + 3;
        } else if (s[i] == '?' || s[i] == '!') {
            const size_t e = s.find('>', i);
            if (e == std::string::npos) fail("unterminated declaration");
            i = e + 1;
        } else if (s[i] == '/') {
            ++i;
            const std::string t = name();
            if (stack.size() < 2 || stack.back()->tag != t) fail("mismatched </" + t + ">");
            stack.pop_back();
            ws();
            if (i >= s.size() || s[i] != '>') fail("expected >");
            ++i;
        } else {
            auto n = std::make_unique<Node>();
            n->tag = name();
            for (;;) {
                ws();
                if (i >= s.size()) fail("unterminated element");
                if (s[i] == '/' || s[i] == '>') break;
                const std::string a = name();
                ws();
                if (i >= s.size() || s[i] != '=') fail("expected =");
                ++i;
                ws();
                if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) fail("expected a quoted value");
                const char q = s[i++];
                const size_t e = s.find(q, i);
                if (e == std::string::npos) fail("unterminated value");
                n->attr[a] = s.substr(i, e - i);
                i = e + 1;
            }
            Node* raw = n.get();
            stack.back()->kids.push_back(std::move(n));
            if (s[i] == '/') {
                ++i;
                if (i >= s.size() || s[i] != '>') fail("expected />");
                ++i;
            } else {
                ++i;
                stack.push_back(raw);
            }
        }
    }
    if (stack.size() != 1) throw std::invalid_argument("URDF: unclosed element <" + stack.back()->tag + ">");
    return root;
}

std::vector<double> numbers(const std::string& s, size_t n, const std::vector<double>& def) {
    if (s.empty()) return def;
    std::istringstream ss(s);
    std::vector<double> v(n);
    for (size_t k = 0; k < n; ++k)
        if (!(ss >> v[k])) throw std::invalid_argument("URDF: expected " + std::to_string(n) + " numbers in '" + s + "'");
    return v;
}

// ------------------------------------------------------------------------------------------ URDF
// 3x4 row-major transforms, composed in double.
struct Xf {
    double m[12];
};
Xf identity() { return Xf{{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0}}; }
Xf mul(const Xf& a, const Xf& b) {
    Xf c;
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 4; ++k) {
            double acc = k == 3 ? a.m[r * 4 + 3] : 0.0;
            for (int j = 0; j < 3; ++j) acc += a.m[r * 4 + j] * b.m[j * 4 + k];
            c.m[r * 4 + k] = acc;
        }
    }
    return c;
}
// Trans(xyz) * Rz(y) Ry(p) Rx(r), the URDF <origin> convention.
Xf origin_of(const Node* o) {
    if (!o) return identity();
    const std::vector<double> xyz = numbers(o->get("xyz"), 3, {0, 0, 0});
    const std::vector<double> rpy = numbers(o->get("rpy"), 3, {0, 0, 0});
    const double cr = std::cos(rpy[0]), sr = std::sin(rpy[0]), cp = std::cos(rpy[1]), sp = std::sin(rpy[1]),
                 cy = std::cos(rpy[2]), sy = std::sin(rpy[2]);
    return Xf{{cy * cp, cy * sp * sr - sy * cr, cy * sp * cr + sy * sr, xyz[0],  //
               sy * cp, sy * sp * sr + cy * cr, sy * sp * cr - cy * sr, xyz[1],  //
               -sp, cp * sr, cp * cr, xyz[2]}};
}

struct Chain {
    std::vector<std::string> link_name;   // n_links = dof + 1
    std::vector<const Node*> link;
    std::vector<Xf> origin;               // per movable joint (fixed joints before it folded in)
    std::vector<std::vector<double>> axis;
    std::vector<double> lower, upper;
};

Chain read_chain(const Node& robot) {
    std::map<std::string, const Node*> links;
    std::multimap<std::string, const Node*> by_parent;
    std::map<std::string, int> is_child;
    for (auto& k : robot.kids) {
        if (k->tag == "link") links[k->get("name")] = k.get();
        if (k->tag == "joint") {
            const Node* p = k->child("parent");
            const Node* c = k->child("child");
            if (!p || !c) throw std::invalid_argument("URDF: joint " + k->get("name") + " without parent/child");
            by_parent.emplace(p->get("link"), k.get());
            is_child[c->get("link")]++;
        }
    }
    std::string rootname;
    for (auto& k : robot.kids)
        if (k->tag == "link" && !is_child.count(k->get("name"))) {
            if (!rootname.empty()) throw std::invalid_argument("URDF: more than one root link");
            rootname = k->get("name");
        }
    if (rootname.empty()) throw std::invalid_argument("URDF: no root link");
    Chain ch;
    ch.link_name.push_back(rootname);
    ch.link.push_back(links[rootname]);
    std::string cur = rootname;
    Xf pending = identity();
    bool pending_fixed = false;
    for (;;) {
        const size_t n = by_parent.count(cur);
        if (n == 0) break;
        if (n > 1) throw std::invalid_argument("URDF: link " + cur + " has several child joints (only serial chains)");
        const Node* j = by_parent.find(cur)->second;
        const std::string type = j->get("type");
        const std::string child = j->child("child")->get("link");
        const Xf o = mul(pending, origin_of(j->child("origin")));
        if (type == "fixed") {
            // folds into the next movable joint; a fixed link with geometry of its own is not supported
            const Node* cl = links.count(child) ? links[child] : nullptr;
            pending = o;
            pending_fixed = cl && cl->child("visual");
            cur = child;
            continue;
        }
        if (type != "revolute" && type != "continuous")
            throw std::invalid_argument("URDF: joint " + j->get("name") + " of type " + type + " (revolute and continuous only)");
        if (pending_fixed) throw std::invalid_argument("URDF: a link with geometry behind a fixed joint before " + j->get("name"));
        std::vector<double> ax = numbers(j->child("axis") ? j->child("axis")->get("xyz") : "", 3, {1, 0, 0});
        const double nrm = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
        if (!(nrm > 0)) throw std::invalid_argument("URDF: zero axis in joint " + j->get("name"));
        for (double& a : ax) a /= nrm;
        double lo = -kPi, hi = kPi;
        if (const Node* lim = j->child("limit")) {
            if (type == "revolute") {
                lo = std::stod(lim->get("lower", "0"));
                hi = std::stod(lim->get("upper", "0"));
            }
        }
        ch.origin.push_back(o);
        ch.axis.push_back(ax);
        ch.lower.push_back(lo);
        ch.upper.push_back(hi);
        ch.link_name.push_back(child);
        ch.link.push_back(links.count(child) ? links[child] : nullptr);
        pending = identity();
        pending_fixed = false;
        cur = child;
    }
    if (ch.origin.empty()) throw std::invalid_argument("URDF: no revolute joint on the chain");
    if (ch.origin.size() > 15) throw std::invalid_argument("URDF: more than 15 movable joints");
    return ch;
}

std::string dir_of(const std::string& p) {
    const size_t s = p.find_last_of('/');
    return s == std::string::npos ? std::string(".") : p.substr(0, s);
}

bool exists(const std::string& p) {
    std::ifstream f(p);
    return (bool) f;
}

// Mesh path of a <mesh filename>: absolute, relative to the URDF's folder, or package://pkg/rest
// looked up as rest (or pkg/rest) from the URDF's folder and its parents.
std::string resolve(const std::string& urdf_dir, const std::string& fn) {
    if (fn.rfind("file://", 0) == 0) return fn.substr(7);
    if (!fn.empty() && fn[0] == '/') return fn;
    if (fn.rfind("package://", 0) == 0) {
        const std::string rest = fn.substr(10);
        const size_t s = rest.find('/');
        const std::string tail = s == std::string::npos ? rest : rest.substr(s + 1);
        std::string d = urdf_dir;
        for (int up = 0; up < 4; ++up) {
            if (exists(d + "/" + tail)) return d + "/" + tail;
            if (exists(d + "/" + rest)) return d + "/" + rest;
            d += "/..";
        }
        return urdf_dir + "/" + tail;
    }
    return urdf_dir + "/" + fn;
}

struct Robot {
    rtcd_urdf u{};
    std::vector<Mesh> mesh;  // per link, link frame
};

Robot read_robot(const std::string& path, bool with_meshes) {
    const std::unique_ptr<Node> doc = parse_xml(read_file(path));
    const Node* robot = doc->child("robot");
    if (!robot) throw std::invalid_argument("URDF: no <robot> element in " + path);
    const Chain ch = read_chain(*robot);
    Robot R;
    rtcd_urdf& u = R.u;
    u.dof = (uint32_t) ch.origin.size();
    u.n_links = u.dof + 1;
    const Xf I = identity();
    for (int k = 0; k < 12; ++k) u.base[k] = (float) I.m[k];
    auto snap = [](double x) { return (float) (std::fabs(x) < 1e-9 ? 0.0 : x); };
    for (uint32_t i = 0; i < u.dof; ++i) {
        const double* R0 = ch.origin[i].m;
        const std::vector<double>& k = ch.axis[i];
        const double K[9] = {0, -k[2], k[1], k[2], 0, -k[0], -k[1], k[0], 0};
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                double cm = 0, sm = 0, om = 0;
                for (int j = 0; j < 3; ++j) {
                    const double kk = k[j] * k[c];
                    cm += R0[r * 4 + j] * ((j == c ? 1.0 : 0.0) - kk);
                    sm += R0[r * 4 + j] * K[j * 3 + c];
                    om += R0[r * 4 + j] * kk;
                }
                u.cos_mask[i][r * 4 + c] = snap(cm);
                u.sin_mask[i][r * 4 + c] = snap(sm);
                u.one_mask[i][r * 4 + c] = snap(om);
            }
            u.cos_mask[i][r * 4 + 3] = 0.f;
            u.sin_mask[i][r * 4 + 3] = 0.f;
            u.one_mask[i][r * 4 + 3] = snap(R0[r * 4 + 3]);
        }
        u.lower[i] = (float) ch.lower[i];
        u.upper[i] = (float) ch.upper[i];
    }
    const std::string dir = dir_of(path);
    for (uint32_t l = 0; l < u.n_links; ++l) {
        std::snprintf(u.link_name[l], sizeof(u.link_name[l]), "%s", ch.link_name[l].c_str());
        const Node* vis = ch.link[l] ? ch.link[l]->child("visual") : nullptr;
        const Node* geo = vis ? vis->child("geometry") : nullptr;
        const Node* mn = geo ? geo->child("mesh") : nullptr;
        if (mn) {
            const std::string file = resolve(dir, mn->get("filename"));
            if (file.size() >= sizeof(u.link_mesh[l])) throw std::invalid_argument("URDF: mesh path too long: " + file);
            std::snprintf(u.link_mesh[l], sizeof(u.link_mesh[l]), "%s", file.c_str());
        }
        const Xf vo = mn ? origin_of(vis->child("origin")) : I;
        const std::vector<double> sc = mn ? numbers(mn->get("scale"), 3, {1, 1, 1}) : std::vector<double>{1, 1, 1};
        for (int k = 0; k < 12; ++k) u.link_mesh_xf[l][k] = (float) vo.m[k];
        for (int k = 0; k < 3; ++k) u.link_mesh_scale[l][k] = (float) sc[k];
        Mesh m;
        if (with_meshes && mn) {
            m = parse_obj(read_file(u.link_mesh[l]), u.link_mesh[l]);
            // scale, then the visual origin, with the float values the URDF record holds
            const float* X = u.link_mesh_xf[l];
            const float* S = u.link_mesh_scale[l];
            const bool moved = std::memcmp(vo.m, I.m, sizeof(vo.m)) != 0 || sc[0] != 1 || sc[1] != 1 || sc[2] != 1;
            if (moved)
                for (size_t v = 0; v < m.v.size(); v += 3) {
                    const double p[3] = {(double) m.v[v] * S[0], (double) m.v[v + 1] * S[1], (double) m.v[v + 2] * S[2]};
                    for (int r = 0; r < 3; ++r)
                        m.v[v + r] = (float) ((double) X[r * 4] * p[0] + (double) X[r * 4 + 1] * p[1] +
                                              (double) X[r * 4 + 2] * p[2] + (double) X[r * 4 + 3]);
                }
        }
        R.mesh.push_back(std::move(m));
    }
    return R;
}

std::vector<float> read_spheres(const std::string& path) {
    std::istringstream in(read_file(path));
    std::string line;
    std::vector<float> s;
    size_t lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        const size_t b = line.find_first_not_of(" \t\r");
        if (b == std::string::npos || line[b] == '#') continue;
        std::istringstream ls(line);
        float x, y, z, r;
        if (!(ls >> x >> y >> z >> r)) throw std::invalid_argument(path + ":" + std::to_string(lineno) + ": expected x y z r");
        if (!(r > 0) || !std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z) || !std::isfinite(r))
            throw std::invalid_argument(path + ":" + std::to_string(lineno) + ": bad sphere");
        s.insert(s.end(), {x, y, z, r});
    }
    return s;
}

template <typename F>
rtcd_status guarded_io(F&& f) {
    try {
        return f();
    } catch (const std::invalid_argument& e) {
        return set_err(RTCD_ERR_INVALID_ARG, e.what());
    } catch (const std::bad_alloc&) {
        return set_err(RTCD_ERR_OOM, "host allocation failed");
    } catch (const std::exception& e) {
        return set_err(RTCD_ERR_INTERNAL, e.what());
    }
}

}  // namespace

extern "C" {

rtcd_status rtcd_obj_read(const char* path, float* verts, uint32_t* n_verts, int32_t* tris, uint32_t* n_tris) {
    if (!path || !n_verts || !n_tris) return set_err(RTCD_ERR_INVALID_ARG, "null argument");
    return guarded_io([&]() -> rtcd_status {
        const Mesh m = parse_obj(read_file(path), path);
        const uint32_t nv = (uint32_t) (m.v.size() / 3), nt = (uint32_t) (m.f.size() / 3);
        if (verts || tris) {
            if (!verts || !tris) return set_err(RTCD_ERR_INVALID_ARG, "give both buffers or neither");
            if (*n_verts < nv || *n_tris < nt) {
                *n_verts = nv;
                *n_tris = nt;
                return set_err(RTCD_ERR_INVALID_ARG, "buffers too small (the counts are returned)");
            }
            std::memcpy(verts, m.v.data(), m.v.size() * sizeof(float));
            std::memcpy(tris, m.f.data(), m.f.size() * sizeof(int32_t));
        }
        *n_verts = nv;
        *n_tris = nt;
        return RTCD_OK;
    });
}

rtcd_status rtcd_mesh_edges(const int32_t* tris, uint32_t n_tris, int32_t mode, int32_t* edges, uint32_t* n_edges) {
    if ((!tris && n_tris) || !n_edges) return set_err(RTCD_ERR_INVALID_ARG, "null argument");
    if (mode != RTCD_EDGES_UNIQUE && mode != RTCD_EDGES_LOOP) return set_err(RTCD_ERR_INVALID_ARG, "bad edge mode");
    return guarded_io([&]() -> rtcd_status {
        std::vector<int32_t> e;
        if (mode == RTCD_EDGES_LOOP) {
            e.reserve((size_t) n_tris * 6);
            for (uint32_t f = 0; f < n_tris; ++f)
                for (int j = 0; j < 3; ++j) e.insert(e.end(), {tris[f * 3 + j], tris[f * 3 + (j + 1) % 3]});
        } else {
            std::unordered_map<uint64_t, uint32_t> seen;
            seen.reserve((size_t) n_tris * 2 + 16);
            for (uint32_t f = 0; f < n_tris; ++f)
                for (int j = 0; j < 3; ++j) {
                    const int32_t a = tris[f * 3 + j], b = tris[f * 3 + (j + 1) % 3];
                    if (a < 0 || b < 0) return set_err(RTCD_ERR_INVALID_ARG, "negative vertex index");
                    const uint64_t key = ((uint64_t) (uint32_t) std::min(a, b) << 32) | (uint32_t) std::max(a, b);
                    if (seen.emplace(key, 0u).second) e.insert(e.end(), {a, b});
                }
        }
        const uint32_t ne = (uint32_t) (e.size() / 2);
        if (edges) {
            if (*n_edges < ne) {
                *n_edges = ne;
                return set_err(RTCD_ERR_INVALID_ARG, "buffer too small (the count is returned)");
            }
            std::memcpy(edges, e.data(), e.size() * sizeof(int32_t));
        }
        *n_edges = ne;
        return RTCD_OK;
    });
}

rtcd_status rtcd_urdf_read(const char* path, rtcd_urdf* out) {
    if (!path || !out) return set_err(RTCD_ERR_INVALID_ARG, "null argument");
    return guarded_io([&]() -> rtcd_status {
        *out = read_robot(path, false).u;
        return RTCD_OK;
    });
}

rtcd_status rtcd_spheres_read(const char* path, float* xyzr, uint32_t* n) {
    if (!path || !n) return set_err(RTCD_ERR_INVALID_ARG, "null argument");
    return guarded_io([&]() -> rtcd_status {
        const std::vector<float> s = read_spheres(path);
        const uint32_t ns = (uint32_t) (s.size() / 4);
        if (xyzr) {
            if (*n < ns) {
                *n = ns;
                return set_err(RTCD_ERR_INVALID_ARG, "buffer too small (the count is returned)");
            }
            std::memcpy(xyzr, s.data(), s.size() * sizeof(float));
        }
        *n = ns;
        return RTCD_OK;
    });
}

rtcd_status rtcd_robot_load_urdf(rtcd_ctx* ctx, const char* urdf_path, const char* sphere_dir) {
    if (!ctx || !urdf_path) return set_err(RTCD_ERR_INVALID_ARG, "null argument");
    Robot R;
    std::vector<float> verts, spheres;
    std::vector<int32_t> tris, sphere_link;
    std::vector<uint32_t> vo, to;
    const rtcd_status st = guarded_io([&]() -> rtcd_status {
        R = read_robot(urdf_path, true);
        vo.push_back(0);
        to.push_back(0);
        for (uint32_t l = 0; l < R.u.n_links; ++l) {
            verts.insert(verts.end(), R.mesh[l].v.begin(), R.mesh[l].v.end());
            tris.insert(tris.end(), R.mesh[l].f.begin(), R.mesh[l].f.end());
            vo.push_back((uint32_t) (verts.size() / 3));
            to.push_back((uint32_t) (tris.size() / 3));
            if (sphere_dir) {
                const std::string f = std::string(sphere_dir) + "/" + R.u.link_name[l] + ".txt";
                if (exists(f)) {
                    const std::vector<float> s = read_spheres(f);
                    spheres.insert(spheres.end(), s.begin(), s.end());
                    sphere_link.insert(sphere_link.end(), s.size() / 4, (int32_t) l);
                }
            }
        }
        return RTCD_OK;
    });
    if (st != RTCD_OK) return st;
    rtcd_robot_desc d{};
    d.dof = R.u.dof;
    d.base = R.u.base;
    d.cos_mask = &R.u.cos_mask[0][0];
    d.sin_mask = &R.u.sin_mask[0][0];
    d.one_mask = &R.u.one_mask[0][0];
    d.verts = verts.data();
    d.link_vert_offsets = vo.data();
    d.tris = tris.data();
    d.link_tri_offsets = to.data();
    d.spheres = spheres.empty() ? nullptr : spheres.data();
    d.sphere_link = sphere_link.empty() ? nullptr : sphere_link.data();
    d.n_spheres = (uint32_t) sphere_link.size();
    return rtcd_robot_load(ctx, &d);
}

}  // extern "C"
