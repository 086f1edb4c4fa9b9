// csg.cpp -- CSG object tree, point inclusion and list assembly (src/cgs.cpp), scene files.
#include <cstdio>
#include <cmath>
#include <fstream>
#include <sstream>

#include "json.hpp"
#include "sjf_host.hpp"

namespace sjfh {

namespace {

Vec3 sub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec3 add(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec3 scale(Vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
Vec3 divide(Vec3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
double norm(Vec3 a) {
    double s = 0;
    s += a.x * a.x;
    s += a.y * a.y;
    s += a.z * a.z;
    return std::sqrt(s);
}
Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// matrix-vector and matrix-matrix products, accumulated left to right from zero
Vec3 mul(const Mat3& m, Vec3 v) {
    double r[3];
    for (int i = 0; i < 3; ++i) {
        double acc = 0;
        acc += m.el[3 * i] * v.x;
        acc += m.el[3 * i + 1] * v.y;
        acc += m.el[3 * i + 2] * v.z;
        r[i] = acc;
    }
    return {r[0], r[1], r[2]};
}
Mat3 mul(const Mat3& a, const Mat3& b) {
    Mat3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0;
            for (int k = 0; k < 3; ++k) acc += a.el[3 * i + k] * b.el[3 * k + j];
            r.el[3 * i + j] = acc;
        }
    return r;
}

int child_in(const Node* c, const Vec3& r, int parent_invert) { return c ? in(*c, r) : parent_invert; }

}  // namespace

Mat3 rotation(double theta, Vec3 a) {
    a = divide(a, norm(a));
    const double ct = std::cos(theta), st = std::sin(theta), ctc = 1.0 - ct;
    Mat3 m;
    m.el[0] = ct + a.x * a.x * ctc;
    m.el[1] = a.x * a.y * ctc - a.z * st;
    m.el[2] = a.x * a.z * ctc + a.y * st;
    m.el[3] = m.el[1] + 2 * a.z * st;
    m.el[4] = ct + a.y * a.y * ctc;
    m.el[5] = a.y * a.z * ctc - a.x * st;
    m.el[6] = m.el[2] - 2 * a.y * st;
    m.el[7] = m.el[5] + 2 * a.x * st;
    m.el[8] = ct + a.z * a.z * ctc;
    return m;
}

// Point inclusion of src/cgs.cpp:34-118 (primitives) and 422-447 (composites).  Primitives
// test trans * (r - centre); a composite transforms r by its own matrix before asking its
// children; planes carry no transform or inversion flag (their inversion is folded into
// the offset when built).
int in(const Node& n, const Vec3& r) {
    switch (n.kind) {
        case Kind::Sphere: {
            const Vec3 q = mul(n.trans, sub(r, n.center));
            return norm(q) > n.radius ? n.invert : 1 - n.invert;
        }
        case Kind::Box: {
            const Vec3 q = mul(n.trans, sub(r, n.center));
            if (std::fabs(q.x) > n.offset.x) return n.invert;
            if (std::fabs(q.y) > n.offset.y) return n.invert;
            if (std::fabs(q.z) > n.offset.z) return n.invert;
            return 1 - n.invert;
        }
        case Kind::Plane:
            return dot(r, n.normal) > n.plane_offset ? 0 : 1;
        case Kind::Cylinder: {
            const Vec3 q = mul(n.trans, sub(r, n.center));
            if (q.z < 0 || q.z > n.height) return n.invert;
            const double rad_sq = q.x * q.x + q.y * q.y;
            if (rad_sq * n.height > (n.r2_sq - n.r1_sq) * q.z + n.r1_sq_x_h) return n.invert;
            return 1 - n.invert;
        }
        case Kind::Composite: {
            const Vec3 q = mul(n.trans, r);
            if (n.combine == Combine::Noop) return 0;
            const Node* l = n.child[0].get();
            const Node* rr = n.child[1].get();
            if (!rr) return l ? (n.invert ^ in(*l, q)) : n.invert;
            if (!l) return n.invert ^ in(*rr, q);
            const int a = child_in(l, q, n.invert), b = child_in(rr, q, n.invert);
            if (n.combine == Combine::Union) return n.invert ^ (a | b);
            return n.invert ^ (a & b);  // intersect; difference by the flipped right child
        }
    }
    return n.invert;
}

// ---- shape descriptions (the .geom constructors, src/cgs_data.cpp:87-300) -------------------

Shape Box(Vec3 c1, Vec3 c2) {
    Shape s;
    s.type = "Box";
    s.a = c1;
    s.b = c2;
    return s;
}
Shape Sphere(Vec3 center, double radius) {
    Shape s;
    s.type = "Sphere";
    s.a = center;
    s.r = radius;
    return s;
}
Shape Cylinder(Vec3 center, double h, double r1, double r2) {
    Shape s;
    s.type = "Cylinder";
    s.a = center;
    s.h = h;
    s.r1 = r1;
    s.r2 = r2;
    return s;
}
Shape Cylinder(Vec3 center, double h, double r1) { return Cylinder(center, h, r1, r1); }
Shape Plane(Vec3 normal, double offset) {
    Shape s;
    s.type = "Plane";
    s.a = normal;
    s.offset = offset;
    return s;
}
Shape Plane(Vec3 p1, Vec3 p2, Vec3 p3) {
    // normal = (p2 - p1) x (p3 - p1), normalised; offset = normal . p1
    Vec3 nrm = cross(sub(p2, p1), sub(p3, p1));
    nrm = divide(nrm, norm(nrm));
    return Plane(nrm, dot(nrm, p1));
}
static Shape list_shape(const char* t, std::vector<Shape> l) {
    Shape s;
    s.type = t;
    s.list = std::move(l);
    return s;
}
Shape Union(std::vector<Shape> l) { return list_shape("Union", std::move(l)); }
Shape Intersect(std::vector<Shape> l) { return list_shape("Intersect", std::move(l)); }
Shape Complement(std::vector<Shape> l) { return list_shape("Complement", std::move(l)); }
Shape Difference(std::vector<Shape> l) {
    if (l.size() < 2) throw std::invalid_argument("Difference needs at least two entries");
    std::vector<Shape> rest(l.begin() + 1, l.end());
    return Intersect({l[0], Complement(std::move(rest))});
}
Shape Rotate(double theta, Vec3 axis, std::vector<Shape> l) {
    Shape s = list_shape("Rotate", std::move(l));
    s.theta = theta;
    s.a = axis;
    return s;
}

namespace {

std::unique_ptr<Node> composite(Combine cmb, int invert) {
    auto n = std::make_unique<Node>();
    n->kind = Kind::Composite;
    n->combine = cmb;
    n->invert = invert;
    return n;
}

void fill_from_list(Node* self, const std::vector<Shape>& list, std::vector<Mat3>& stack);

// make_object (src/cgs.cpp:122-196): primitives and Union/Intersect nodes, or nullptr for the
// Complement/Rotate markers handled by the caller.
std::unique_ptr<Node> make_object(const Shape& s, int invert, std::vector<Mat3>& stack) {
    auto n = std::make_unique<Node>();
    n->invert = invert;
    if (s.type == "Box") {
        n->kind = Kind::Box;
        Vec3 off = divide(sub(s.b, s.a), 2);
        n->center = add(s.a, off);
        if (off.x < 0) off.x *= -1;
        if (off.y < 0) off.y *= -1;
        if (off.z < 0) off.z *= -1;
        n->offset = off;
    } else if (s.type == "Sphere") {
        n->kind = Kind::Sphere;
        n->center = s.a;
        n->radius = s.r;
    } else if (s.type == "Cylinder") {
        n->kind = Kind::Cylinder;
        n->center = s.a;
        n->height = s.h;
        n->r1_sq = s.r1 * s.r1;
        n->r1_sq_x_h = n->r1_sq * s.h;
        n->r2_sq = s.r2 * s.r2;
    } else if (s.type == "Plane") {
        n->kind = Kind::Plane;
        n->invert = 0;
        n->plane_offset = invert ? -s.offset : s.offset;
        n->normal = divide(s.a, norm(s.a));
    } else if (s.type == "Union" || s.type == "Intersect") {
        n = composite(s.type == "Union" ? Combine::Union : Combine::Intersect, invert);
        fill_from_list(n.get(), s.list, stack);
    } else {
        return nullptr;
    }
    return n;
}

// composite_object::append (src/cgs.cpp:200-212): fill the left slot, the right slot for the
// last entry, otherwise chain a new node of the same combination into the right slot.
void append(Node** last, std::unique_ptr<Node> obj, bool is_last) {
    Node* lc = *last;
    if (!lc->child[0]) {
        lc->child[0] = std::move(obj);
    } else if (is_last) {
        lc->child[1] = std::move(obj);
    } else {
        auto nc = composite(lc->combine, 0);
        Node* raw = nc.get();
        lc->child[1] = std::move(nc);
        *last = raw;
        raw->child[0] = std::move(obj);
    }
}

// composite_object::init_from_list (src/cgs.cpp:214-270).  The chained nodes appended to a list
// share the combination of the list's owner; a Complement marker flips the inversion used for
// every later entry and becomes a node whose own flag is the inversion before the flip;
// Rotate pushes its matrix on the shared stack (never popped) and becomes a node built with
// the grown stack.
void fill_from_list(Node* self, const std::vector<Shape>& list, std::vector<Mat3>& stack) {
    Node* last = self;
    int invert = 0;
    Mat3 trans;
    for (size_t i = stack.size(); i > 0; --i) trans = mul(trans, stack[i - 1]);
    const size_t n = list.size();
    for (size_t i = 0; i < n; ++i) {
        const Shape& s = list[i];
        const bool is_last = (i == n - 1);
        if (s.type.empty()) continue;  // a list entry that is no object: skipped, but counted
        std::unique_ptr<Node> obj = make_object(s, invert, stack);
        if (!obj) {
            if (s.type == "Complement") {
                invert = 1 - invert;
                auto c = composite(self->combine, 1 - invert);
                fill_from_list(c.get(), s.list, stack);
                append(&last, std::move(c), is_last);
            } else if (s.type == "Rotate") {
                stack.push_back(rotation(s.theta, s.a));
                auto c = composite(self->combine, invert);
                fill_from_list(c.get(), s.list, stack);
                append(&last, std::move(c), is_last);
            } else {
                break;  // unknown entry: the reference stops reading the list
            }
        } else {
            obj->trans = trans;
            append(&last, std::move(obj), is_last);
        }
    }
}

}  // namespace

std::unique_ptr<Node> assemble(Combine cmb, const std::vector<Shape>& list) {
    auto root = composite(cmb, 0);
    std::vector<Mat3> stack;
    fill_from_list(root.get(), list, stack);
    return root;
}

// ---- scene files ---------------------------------------------------------------------------

namespace {

Vec3 vec_of(const json::J& j) { return {j[0].num(), j[1].num(), j[2].num()}; }

std::unique_ptr<Node> node_of(const json::J& j) {
    if (j.type == json::J::Null) return nullptr;
    auto n = std::make_unique<Node>();
    const std::string k = j.at("kind").s;
    n->invert = (int)j.at("invert").num();
    if (j.has("trans"))
        for (int q = 0; q < 9; ++q) n->trans.el[q] = j.at("trans")[q].num();
    if (k == "composite") {
        n->kind = Kind::Composite;
        const std::string c = j.at("combine").s;
        n->combine = c == "union" ? Combine::Union
                   : c == "intersect" ? Combine::Intersect
                   : c == "difference" ? Combine::Difference : Combine::Noop;
        n->child[0] = node_of(j.at("children")[0]);
        n->child[1] = node_of(j.at("children")[1]);
    } else if (k == "sphere") {
        n->kind = Kind::Sphere;
        n->center = vec_of(j.at("center"));
        n->radius = j.at("radius").num();
    } else if (k == "box") {
        n->kind = Kind::Box;
        n->center = vec_of(j.at("center"));
        n->offset = vec_of(j.at("offset"));
    } else if (k == "plane") {
        n->kind = Kind::Plane;
        n->normal = vec_of(j.at("normal"));
        n->plane_offset = j.at("offset").num();
    } else if (k == "cylinder") {
        n->kind = Kind::Cylinder;
        n->center = vec_of(j.at("center"));
        n->height = j.at("height").num();
        n->r1_sq = j.at("r1_sq").num();
        n->r1_sq_x_h = j.at("r1_sq_x_h").num();
        n->r2_sq = j.at("r2_sq").num();
    } else {
        throw std::runtime_error("scene: unknown node kind '" + k + "'");
    }
    return n;
}

Value value_of(const json::J& j) {
    Value v;
    switch (j.type) {
        case json::J::Num: v.type = Value::Num; v.x = j.x; break;
        case json::J::Str: v.type = Value::Str; v.s = j.s; break;
        case json::J::Arr:
            v.type = Value::List;
            for (const auto& e : j.a) v.l.push_back(value_of(e));
            break;
        default: break;
    }
    return v;
}

}  // namespace

Scene Scene::from_json_text(const std::string& text) {
    const json::J d = json::parse(text);
    Scene sc;
    if (d.has("roots")) {
        for (const auto& r : d.at("roots").a) {
            Root root;
            if (r.has("metadata"))
                for (const auto& kv : r.at("metadata").o) root.metadata[kv.first] = value_of(kv.second);
            root.tree = node_of(r.at("tree"));
            sc.roots.push_back(std::move(root));
        }
    }
    if (d.has("sources")) {
        static const char* keys[9] = {"component", "wavelength", "amplitude", "width", "phase",
                                      "cutoff", "start_time", "end_time", "slowness"};
        for (const auto& s : d.at("sources").a) {
            SourceSpec sp;
            sp.gaussian = s.at("type").s == "Gaussian_source";
            double* dst[9] = {&sp.component, &sp.wavelength, &sp.amplitude, &sp.width, &sp.phase,
                              &sp.cutoff, &sp.start_time, &sp.end_time, &sp.slowness};
            for (int q = 0; q < 9; ++q)
                if (s.has(keys[q])) {
                    sp.has[q] = true;
                    *dst[q] = s.at(keys[q]).num();
                }
            if (s.has("box")) {
                sp.has_box = true;
                sp.box[0] = vec_of(s.at("box")[0]);
                sp.box[1] = vec_of(s.at("box")[1]);
            }
            sc.sources.push_back(sp);
        }
    }
    if (d.has("monitor_clusters")) {
        for (const auto& c : d.at("monitor_clusters").a) {
            std::vector<Vec3> pts;
            for (const auto& p : c.a) pts.push_back(vec_of(p));
            sc.monitor_clusters.push_back(std::move(pts));
        }
    }
    if (d.has("context_numbers"))
        for (const auto& kv : d.at("context_numbers").a) sc.context_numbers.emplace_back(kv[0].s, kv[1].num());
    return sc;
}

namespace {

std::string jnum(double x) {
    char b[40];
    std::snprintf(b, sizeof b, "%.17g", x);
    return b;
}

std::string jstr(const std::string& s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        if ((unsigned char)c < 0x20) {
            o += ' ';
            continue;
        }
        o += c;
    }
    return o + "\"";
}

std::string jvec(const Vec3& v) { return "[" + jnum(v.x) + "," + jnum(v.y) + "," + jnum(v.z) + "]"; }

std::string jvalue(const Value& v) {
    switch (v.type) {
        case Value::Num: return jnum(v.x);
        case Value::Str: return jstr(v.s);
        case Value::List: {
            std::string o = "[";
            for (size_t i = 0; i < v.l.size(); ++i) o += (i ? "," : "") + jvalue(v.l[i]);
            return o + "]";
        }
        default: return "null";
    }
}

std::string jnode(const Node* n) {
    if (!n) return "null";
    static const char* kinds[5] = {"composite", "sphere", "box", "plane", "cylinder"};
    std::string o = "{\"kind\":\"" + std::string(kinds[(int)n->kind]) + "\",\"invert\":" + std::to_string(n->invert) +
                    ",\"trans\":[";
    for (int q = 0; q < 9; ++q) o += (q ? "," : "") + jnum(n->trans.el[q]);
    o += "]";
    switch (n->kind) {
        case Kind::Composite: {
            const char* c = n->combine == Combine::Union ? "union" : n->combine == Combine::Intersect ? "intersect"
                          : n->combine == Combine::Difference ? "difference" : "noop";
            o += ",\"combine\":\"" + std::string(c) + "\",\"children\":[" + jnode(n->child[0].get()) + "," +
                 jnode(n->child[1].get()) + "]";
            break;
        }
        case Kind::Sphere: o += ",\"center\":" + jvec(n->center) + ",\"radius\":" + jnum(n->radius); break;
        case Kind::Box: o += ",\"center\":" + jvec(n->center) + ",\"offset\":" + jvec(n->offset); break;
        case Kind::Plane: o += ",\"normal\":" + jvec(n->normal) + ",\"offset\":" + jnum(n->plane_offset); break;
        case Kind::Cylinder:
            o += ",\"center\":" + jvec(n->center) + ",\"height\":" + jnum(n->height) + ",\"r1_sq\":" + jnum(n->r1_sq) +
                 ",\"r1_sq_x_h\":" + jnum(n->r1_sq_x_h) + ",\"r2_sq\":" + jnum(n->r2_sq);
            break;
    }
    return o + "}";
}

}  // namespace

std::string Scene::to_json() const {
    std::string o = "{\"roots\":[";
    for (size_t r = 0; r < roots.size(); ++r) {
        o += std::string(r ? "," : "") + "{\"metadata\":{";
        bool first = true;
        for (const auto& kv : roots[r].metadata) {
            o += (first ? "" : ",") + jstr(kv.first) + ":" + jvalue(kv.second);
            first = false;
        }
        o += "},\"tree\":" + jnode(roots[r].tree.get()) + "}";
    }
    o += "],\"sources\":[";
    static const char* keys[9] = {"component", "wavelength", "amplitude", "width", "phase",
                                  "cutoff", "start_time", "end_time", "slowness"};
    for (size_t i = 0; i < sources.size(); ++i) {
        const SourceSpec& sp = sources[i];
        const double* val[9] = {&sp.component, &sp.wavelength, &sp.amplitude, &sp.width, &sp.phase,
                                &sp.cutoff, &sp.start_time, &sp.end_time, &sp.slowness};
        o += std::string(i ? "," : "") + "{\"type\":" + jstr(sp.gaussian ? "Gaussian_source" : "CW_source");
        for (int q = 0; q < 9; ++q)
            if (sp.has[q]) o += "," + jstr(keys[q]) + ":" + jnum(*val[q]);
        if (sp.has_box) o += ",\"box\":[" + jvec(sp.box[0]) + "," + jvec(sp.box[1]) + "]";
        o += "}";
    }
    o += "],\"monitor_clusters\":[";
    for (size_t k = 0; k < monitor_clusters.size(); ++k) {
        o += std::string(k ? "," : "") + "[";
        for (size_t p = 0; p < monitor_clusters[k].size(); ++p) o += (p ? "," : "") + jvec(monitor_clusters[k][p]);
        o += "]";
    }
    o += "],\"context_numbers\":[";
    for (size_t k = 0; k < context_numbers.size(); ++k)
        o += std::string(k ? "," : "") + "[" + jstr(context_numbers[k].first) + "," + jnum(context_numbers[k].second) + "]";
    return o + "]}";
}

Scene Scene::from_json_file(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open scene file " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return from_json_text(ss.str());
}

}  // namespace sjfh
