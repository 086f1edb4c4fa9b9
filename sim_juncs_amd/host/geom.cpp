// geom.cpp -- reader of the .geom scene language (the subset the reference's junctions and
// tests use), producing the Scene that voxelize() and BoundGeom consume.
//
// The language is the reference's own (src/cgs_read.cpp, src/cgs_data.cpp, src/cgs.cpp:510-
// 570, context set-up src/disp.cpp:5-50).  Its evaluator is string-splitting rather than
// grammar-based, and scenes written for it depend on the results that gives, so this reader
// keeps the same evaluation rules, quirks included:
//   * an expression splits at the first operator, outside brackets and quotes, of the loosest
//     class present: comparisons (==, <, >, <=, >=), then + and -, then ^, then * and /, then
//     the ternary ?: -- so a - b - c is a - (b - c), and x + c ? a : b is x + (c ? a : b);
//     a + or - right after an 'e' never splits (exponents, but also names ending in e);
//   * an undefined name evaluates to the number strtod gives for it (0 for a name);
//   * list and argument elements lose all whitespace outside brackets and quotes, so
//     "[0.33 0.19 11.4, "lorentz"]" is the two-element list [0.33, "lorentz"] (strtod stops at
//     the second '.' of the fused token);
//   * a trailing comma adds a last element that evaluates to 0; in a geometry list such an
//     element is skipped but still counts when the tree decides which entry is last;
//   * an error inside a function-call argument other than the last one is dropped;
//   * statements end at a newline or ';', except that a bracket or quote opened on a line runs
//     on to its closing partner; a statement whose block closes on a later line ends there.
// Not supported: user function definitions (def) and member access (a.b), which no junction or
// test scene uses; they report an error.  snapshot() makes its instance but renders no image.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <limits>
#include <memory>
#include <sstream>

#include "sjf_host.hpp"

namespace sjfh {
namespace {

// the reference's parse_ercode values (src/cgs_read.hpp:32)
enum Err {
    OK = 0,
    E_NOFILE = 1,
    E_LACK_TOKENS = 2,
    E_BAD_TOKEN = 3,
    E_BAD_SYNTAX = 4,
    E_BAD_VALUE = 5,
    E_BAD_TYPE = 6,
    E_NOMEM = 7,
    E_NAN = 10,
    E_NOT_DEFINED = 11,
    E_OUT_OF_RANGE = 12
};

constexpr size_t kMaxArgs = 256;  // argument buffer of a call (src/cgs_read.hpp:17)
constexpr size_t kMaxStr = 1024;  // string concatenation buffer (src/cgs_read.hpp:14)

enum Builtin {
    F_NONE,
    F_UM_TO_L,
    F_FS_TO_T,
    F_GAUSSIAN,
    F_CW,
    F_BOX,
    F_PLANE,
    F_SPHERE,
    F_CYLINDER,
    F_COMPOSITE,
    F_UNION,
    F_INTERSECT,
    F_COMPLEMENT,
    F_ROTATE,
    F_DIFFERENCE,
    F_SNAPSHOT,
    F_MONITORS
};

struct GV;
using Fields = std::vector<std::pair<std::string, GV>>;

// a value of the language: number, string, list, 3-vector, function or instance (an object made
// by a constructor such as Box(): its type name and named fields, oldest first)
struct GV {
    enum T { Undef, Str, Num, List, Vec, Func, Inst } t = Undef;
    double x = 0;
    std::string s;
    std::vector<GV> l;
    Vec3 v;
    int fn = F_NONE, min_args = 0;
    std::shared_ptr<Fields> f;
    long lit = -1;  // instance written as a literal {a, n=b, ...}: its element count

    static GV num(double d) {
        GV g;
        g.t = Num;
        g.x = d;
        return g;
    }
    static GV str(std::string x) {
        GV g;
        g.t = Str;
        g.s = std::move(x);
        return g;
    }
    static GV vec(Vec3 x) {
        GV g;
        g.t = Vec;
        g.v = x;
        return g;
    }
    static GV inst(const char* type) {
        GV g;
        g.t = Inst;
        g.f = std::make_shared<Fields>();
        g.f->emplace_back("__type__", str(type));
        return g;
    }
    void put(const std::string& name, const GV& val) { f->emplace_back(name, val); }
    // newest field of that name
    GV get(const std::string& name) const {
        if (t != Inst) return GV();
        for (size_t i = f->size(); i > 0; --i)
            if ((*f)[i - 1].first == name) return (*f)[i - 1].second;
        return GV();
    }
    void set(const std::string& name, const GV& val) {
        for (size_t i = f->size(); i > 0; --i)
            if ((*f)[i - 1].first == name) {
                (*f)[i - 1].second = val;
                return;
            }
        put(name, val);
    }
    std::string type_name() const {
        const GV ty = get("__type__");
        return ty.t == Str ? ty.s : std::string();
    }
};

// list -> 3-vector (at least three numbers), as value::cast_to (src/cgs_read.cpp:1206-1227)
int to_vec(const GV& g, Vec3* out) {
    if (g.t == GV::Vec) {
        *out = g.v;
        return OK;
    }
    if (g.t == GV::Undef) return E_BAD_VALUE;
    if (g.t != GV::List) return E_BAD_VALUE;
    if (g.l.size() < 3) return E_LACK_TOKENS;
    if (g.l[0].t != GV::Num || g.l[1].t != GV::Num || g.l[2].t != GV::Num) return E_BAD_TOKEN;
    *out = {g.l[0].x, g.l[1].x, g.l[2].x};
    return OK;
}

// 3-vector or instance -> list (value::cast_to, src/cgs_read.cpp:1228-1253).  An instance
// literal of n entries gives n elements: an undefined one, then its entries newest first
// (the oldest does not fit); a constructed instance gives an empty list.
GV to_list(const GV& g, int* er) {
    *er = OK;
    if (g.t == GV::List) return g;
    if (g.t == GV::Vec) {
        GV r;
        r.t = GV::List;
        r.l = {GV::num(g.v.x), GV::num(g.v.y), GV::num(g.v.z)};
        return r;
    }
    if (g.t == GV::Inst) {
        GV r;
        r.t = GV::List;
        const size_t size = g.f->size(), n = g.lit > 0 ? (size_t)g.lit : 0;
        r.l.resize(n);
        for (size_t i = 1; i < size && i < n; ++i) r.l[i] = (*g.f)[size - i].second;
        return r;
    }
    *er = g.t == GV::Undef ? E_BAD_VALUE : E_BAD_TYPE;
    return GV();
}

bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n'; }

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && is_ws(s[a])) ++a;
    while (b > a && is_ws(s[b - 1])) --b;
    return s.substr(a, b - a);
}

// separators that end a token (is_char_sep, src/cgs_read.cpp:8-14)
bool is_sep(char c) {
    return c == 0 || c == ' ' || c == '\t' || c == '\n' || c == ';' || c == '+' || c == '-' || c == '*' || c == '/';
}

bool is_open(char c) { return c == '(' || c == '[' || c == '{'; }
bool is_close(char c) { return c == ')' || c == ']' || c == '}'; }
char partner(char c) { return c == '(' ? ')' : c == '[' ? ']' : c == '{' ? '}' : c; }

// first index of c outside brackets and quotes (strchr_block, src/cgs_read.cpp:38-74), or npos
size_t find_outside(const std::string& s, size_t from, char c) {
    std::vector<char> st;
    for (size_t i = from; i < s.size(); ++i) {
        const char ch = s[i];
        if (ch == c && st.empty()) return i;
        if (is_open(ch)) {
            st.push_back(ch);
        } else if (is_close(ch)) {
            if (!st.empty()) st.pop_back();
        } else if (ch == '"' || ch == '\'') {
            if (!st.empty() && st.back() == ch) st.pop_back();
            else st.push_back(ch);
        }
    }
    return std::string::npos;
}

// first occurrence of the word w outside brackets, quotes and comments, delimited by token
// separators (token_block, src/cgs_read.cpp:76-126), or npos
size_t find_word(const std::string& s, size_t from, const char* w) {
    const size_t n = std::strlen(w);
    std::vector<char> st;
    for (size_t i = from; i < s.size(); ++i) {
        if (st.empty() && s.compare(i, n, w) == 0) {
            const bool before = i == 0 || is_sep(s[i - 1]);
            const bool after = i + n >= s.size() || is_sep(s[i + n]);
            if (before && after) return i;
        }
        const char ch = s[i];
        if (is_open(ch)) {
            st.push_back(ch);
        } else if (is_close(ch)) {
            if (!st.empty()) st.pop_back();
        } else if (i > 0 && s[i - 1] == '/' && (ch == '*' || ch == '/')) {
            st.push_back('/');
        } else if (i > 0 && s[i - 1] == '*' && ch == '/') {
            if (!st.empty()) st.pop_back();
        } else if (ch == '"' || ch == '\'') {
            if (!st.empty() && st.back() == ch) st.pop_back();
            else st.push_back(ch);
        }
    }
    return std::string::npos;
}

// split at sep outside brackets and quotes, dropping whitespace outside them and resolving
// escapes (csv_to_list, src/cgs_read.cpp:211-301).  An element is added after the last separator
// whenever anything at all was kept, so a trailing separator yields an empty last element.
std::vector<std::string> split_elems(const std::string& s, char sep, int* er) {
    *er = OK;
    std::vector<std::string> out;
    std::vector<char> st;
    std::string cur;
    bool wrote = false, verbatim = false;
    for (size_t i = 0; i < s.size(); ++i) {
        char ch = s[i];
        if (ch == sep && st.empty()) {
            out.push_back(cur);
            cur.clear();
            wrote = true;
            continue;
        }
        if (ch == '\\') {
            ++i;
            const char e = i < s.size() ? s[i] : 0;
            const char put = e == 'n' ? '\n' : e == 't' ? '\t' : e == '\\' ? '\\' : e == '"' ? '"' : 0;
            if (put) {
                cur += put;
                wrote = true;
                continue;
            }
            *er = E_BAD_SYNTAX;
            if (i >= s.size()) break;
            ch = s[i];
        } else if (ch == '"') {
            if (st.empty() || st.back() != '"') {
                st.push_back('"');
                verbatim = true;
            } else {
                st.pop_back();
                verbatim = false;
            }
        } else if (is_open(ch)) {
            st.push_back(ch);
        } else if (is_close(ch)) {
            if (st.empty()) break;  // the end of the enclosing list
            if (partner(st.back()) != ch) {
                *er = E_BAD_SYNTAX;
                return {};
            }
            st.pop_back();
        }
        if (!st.empty() || verbatim || !is_ws(ch)) {
            cur += ch;
            wrote = true;
        }
    }
    if (!st.empty()) {
        *er = E_BAD_SYNTAX;
        return {};
    }
    if (wrote) out.push_back(cur);
    return out;
}

struct Entry {
    bool named;
    std::string name;
    GV v;
};

struct Call {
    std::string name;
    std::vector<GV> args;
    std::vector<std::string> names;  // "" for positional arguments
    std::vector<bool> has_name;
    GV named(const char* n) const {
        for (size_t i = 0; i < args.size(); ++i)
            if (has_name[i] && names[i] == n) return args[i];
        return GV();
    }
};

class Interp {
public:
    std::vector<Entry> st;  // context entries, oldest first
    int er = OK;

    GV lookup(const std::string& name) const {
        for (size_t i = st.size(); i > 0; --i)
            if (st[i - 1].named && st[i - 1].name == name) return st[i - 1].v;
        return GV();
    }
    void emplace(const std::string* name, const GV& v) { st.push_back({name != nullptr, name ? *name : "", v}); }
    void set_value(const std::string& name, const GV& v) {
        for (size_t i = st.size(); i > 0; --i)
            if (st[i - 1].named && st[i - 1].name == name) {
                st[i - 1].v = v;
                return;
            }
        emplace(&name, v);
    }
    void define(const char* name, const GV& v) {
        const std::string n(name);
        emplace(&n, v);
    }
    void define_func(const char* name, int fn, int min_args) {
        GV g;
        g.t = GV::Func;
        g.fn = fn;
        g.min_args = min_args;
        define(name, g);
    }

    GV eval(const std::string& in);
    int read_lines(const std::vector<std::string>& lines);

private:
    GV do_op(const std::string& s, size_t i);
    GV parse_list(const std::string& s);
    GV call(const std::string& s, size_t open, size_t close);
    GV construct(const Call& c);
    int read_statement(const std::vector<std::string>& lines, size_t* line, size_t* off);
};

// parse_value (src/cgs_read.cpp:1780-2088), restated: find the splitting operator at nesting
// level 0; without one, the text is a name or number, a string, a list or an index into one, a
// call or a parenthesised expression.
GV Interp::eval(const std::string& in) {
    er = OK;
    size_t b = 0;
    while (b < in.size() && is_ws(in[b])) ++b;
    const std::string s = in.substr(b);
    const bool numeric = !s.empty() && (s[0] == '+' || s[0] == '-' || s[0] == '.' || (s[0] >= '0' && s[0] <= '9'));
    long first_open = -1, last_close = -1;
    int prec = 0;
    size_t op = 0;
    int nest = 0;
    std::vector<char> blk;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        const char nx = i + 1 < s.size() ? s[i + 1] : 0;
        if (is_open(c)) {
            blk.push_back(c);
            ++nest;
            if (first_open < 0) first_open = (long)i;
        } else if (is_close(c)) {
            if (blk.empty() || partner(blk.back()) != c) {
                er = E_BAD_SYNTAX;
                return GV();
            }
            blk.pop_back();
            --nest;
            last_close = (long)i;
        } else if (c == '"' && (i == 0 || s[i - 1] != '\\')) {
            if (blk.empty() || blk.back() != '"') {
                blk.push_back('"');
                ++nest;
                if (first_open < 0) first_open = (long)i;
            } else {
                blk.pop_back();
                last_close = (long)i;
                --nest;
            }
        } else if (c == '/' && nx == '/') {
            break;
        }
        if (nest == 0) {
            if (((c == '=' && nx == '=') || c == '>' || c == '<') && prec < 5) {
                prec = 5;
                op = i;
            } else if (i != 0 && (c == '+' || c == '-') && s[i - 1] != 'e' && prec < 4) {
                prec = 4;
                op = i;
            } else if (c == '^' && prec < 3) {
                prec = 3;
                op = i;
            } else if ((c == '*' || c == '/') && prec < 2) {
                prec = 2;
                op = i;
            } else if (prec < 1 && (c == '?' || (c == '.' && !numeric))) {
                prec = 1;
                op = i;
            }
        }
    }
    if (nest > 0) {
        er = E_BAD_SYNTAX;
        return GV();
    }
    if (prec > 0) return do_op(s, op);

    if (first_open < 0 || last_close < 0) {
        const std::string tok = trim(s);
        GV v = lookup(tok);
        if (v.t != GV::Undef) return v;
        if (tok == "false") return GV::num(0);
        if (tok == "true") return GV::num(1);
        errno = 0;
        const double x = std::strtod(tok.c_str(), nullptr);
        if (errno) {
            er = E_NOT_DEFINED;
            return GV();
        }
        return GV::num(x);
    }
    const char o = s[first_open], c = s[last_close];
    if (o == '"' && c == '"') return GV::str(s.substr(first_open + 1, last_close - first_open - 1));
    if (o == '[' && c == ']') {
        // the token before the bracket names a list to index; none: a list literal
        size_t t = (size_t)first_open;
        while (t > 0 && !is_sep(s[t - 1])) --t;
        const std::string name = s.substr(t, first_open - t);
        if (name.empty()) return parse_list(s.substr(first_open));
        const GV lst = lookup(name);
        if (lst.t != GV::List) {
            er = E_BAD_TYPE;
            return GV();
        }
        const GV idx = eval(s.substr(first_open + 1, last_close - first_open - 1));
        if (er != OK) return GV();
        if (idx.t != GV::Num) {
            er = E_BAD_TYPE;
            return GV();
        }
        long k = (long)idx.x;
        if (k < 0) k += (long)lst.l.size();
        if (k < 0 || (size_t)k >= lst.l.size()) {
            er = E_OUT_OF_RANGE;
            return GV();
        }
        return lst.l[k];
    }
    if (o == '(' && c == ')') {
        for (long j = 0; j < first_open; ++j)
            if (!is_ws(s[j])) return call(s, (size_t)first_open, (size_t)last_close);
        return eval(trim(s.substr(first_open + 1, last_close - first_open - 1)));
    }
    if (o == '{' && c == '}') {
        // an instance literal: entries "v" or "name=v" separated by commas, each evaluated with
        // the earlier ones visible (a child context of this one)
        const std::vector<std::string> els = split_elems(s.substr(first_open + 1, last_close - first_open - 1), ',', &er);
        if (er != OK) return GV();
        const size_t base = st.size();
        for (const std::string& e : els) {
            const size_t eq = find_outside(e, 0, '=');
            const std::string name = eq == std::string::npos ? std::string() : trim(e.substr(0, eq));
            const GV v = eval(eq == std::string::npos ? e : e.substr(eq + 1));
            if (er != OK) {
                st.resize(base);
                return GV();
            }
            emplace(eq == std::string::npos ? nullptr : &name, v);
        }
        GV r;
        r.t = GV::Inst;
        r.f = std::make_shared<Fields>();
        r.lit = (long)els.size();
        for (size_t i = base; i < st.size(); ++i) r.f->emplace_back(st[i].name, st[i].v);
        st.resize(base);
        return r;
    }
    return GV();
}

// do_op (src/cgs_read.cpp:1646-1778): the ternary splits at the first ':' after the '?' outside
// brackets; binary operators evaluate both sides and combine numbers (strings concatenate
// under +).
GV Interp::do_op(const std::string& s, size_t i) {
    const char c = s[i];
    const size_t width = (i + 1 < s.size() && s[i + 1] == '=') ? 2 : 1;
    if (c == '?') {
        const size_t col = find_outside(s, i + 1, ':');
        if (col == std::string::npos) {
            er = E_BAD_SYNTAX;
            return GV();
        }
        const GV cond = eval(s.substr(0, i));
        if (er != OK) return GV();
        if (cond.t == GV::Undef || cond.x == 0) return eval(s.substr(col + 1));
        return eval(s.substr(i + width, col - i - width));
    }
    if (c == '.') {
        std::fprintf(stderr, "geom: member access is not supported\n");
        er = E_BAD_TYPE;
        return GV();
    }
    const GV l = eval(s.substr(0, i));
    if (er != OK) return GV();
    const GV r = eval(s.substr(i + width));
    if (er != OK) return GV();
    GV out = GV::num(0);
    const bool nn = l.t == GV::Num && r.t == GV::Num;
    if (c == '=') {
        bool eq = l.t == r.t;
        if (eq && l.t == GV::Num) eq = l.x == r.x;
        else if (eq && l.t == GV::Str) eq = l.s == r.s;
        else if (eq && l.t != GV::List) eq = false;
        else if (eq) {
            // element-wise (numbers and strings)
            eq = l.l.size() == r.l.size();
            for (size_t k = 0; eq && k < l.l.size(); ++k) {
                const GV &a = l.l[k], &bb = r.l[k];
                eq = a.t == bb.t && ((a.t == GV::Num && a.x == bb.x) || (a.t == GV::Str && a.s == bb.s));
            }
        }
        out.x = eq ? 1 : 0;
    } else if (c == '>' || c == '<') {
        if (!nn) {
            er = E_BAD_VALUE;
            return GV();
        }
        const bool ge = width == 2;
        out.x = c == '>' ? (ge ? l.x >= r.x : l.x > r.x) : (ge ? l.x <= r.x : l.x < r.x);
    } else if (c == '+') {
        if (nn) {
            out.x = l.x + r.x;
        } else if (l.t == GV::Str || r.t == GV::Str) {
            auto rep = [](const GV& g) -> std::string {
                char b[64];
                if (g.t == GV::Str) return g.s;
                if (g.t == GV::Num) {
                    std::snprintf(b, sizeof b, "%f", g.x);
                    return b;
                }
                if (g.t == GV::Vec) {
                    std::snprintf(b, sizeof b, "(%f, %f, %f)", g.v.x, g.v.y, g.v.z);
                    return b;
                }
                return std::string();
            };
            std::string cat = rep(l) + rep(r);
            if (cat.size() > kMaxStr - 1) cat.resize(kMaxStr - 1);
            out = GV::str(cat);
        }
    } else if (c == '-') {
        if (nn) out.x = l.x - r.x;
    } else if (c == '*') {
        if (nn) out.x = l.x * r.x;
    } else if (c == '/') {
        if (r.x == 0) {
            er = E_NAN;
            return GV();
        }
        if (nn) out.x = l.x / r.x;
    } else if (c == '^') {
        if (!nn) {
            er = E_NAN;
            return GV();
        }
        out.x = std::pow(l.x, r.x);
    }
    return out;
}

// parse_list (src/cgs_read.cpp:1570-1641): "[e for v in list]" evaluates e for each element with
// v bound in the context; otherwise the elements separated by commas.
GV Interp::parse_list(const std::string& s) {
    const size_t start = s.find('[');
    const size_t end = find_outside(s, start + 1, ']');
    if (start == std::string::npos || end == std::string::npos) {
        er = E_BAD_SYNTAX;
        return GV();
    }
    const std::string body = s.substr(start + 1, end - start - 1);
    GV out;
    out.t = GV::List;
    const size_t fpos = find_word(body, 0, "for");
    if (fpos != std::string::npos) {
        const size_t ipos = find_word(body, fpos + 3, "in");
        if (ipos == std::string::npos) {
            er = E_BAD_SYNTAX;
            return GV();
        }
        const std::string var = trim(body.substr(fpos + 3, ipos - fpos - 3));
        const std::string expr = body.substr(0, fpos);
        const GV it = eval(body.substr(ipos + 2));
        if (er != OK || it.t != GV::List) {
            er = E_BAD_VALUE;
            return GV();
        }
        emplace(&var, GV());
        for (const GV& e : it.l) {
            set_value(var, e);
            out.l.push_back(eval(expr));
            if (er != OK) return GV();
        }
        st.pop_back();
    } else {
        const std::vector<std::string> els = split_elems(body, ',', &er);
        if (er != OK) return GV();
        for (const std::string& e : els) {
            out.l.push_back(eval(e));
            if (er != OK) return GV();
        }
    }
    er = OK;
    return out;
}

// A call "name(args)": built-in functions by name first, then functions defined in the context
// (src/cgs_read.cpp:1431-1504, 2004-2071).  Arguments split like list elements; "n=v" names one.
GV Interp::call(const std::string& s, size_t open, size_t close) {
    Call c;
    c.name = trim(s.substr(0, open));
    const std::string tail = s.substr(0, close + 1);
    const size_t term = find_outside(tail, open + 1, ')');
    if (term == std::string::npos) {
        er = E_BAD_SYNTAX;
        return GV();
    }
    const std::vector<std::string> els = split_elems(tail.substr(open + 1, term - open - 1), ',', &er);
    if (er != OK) return GV();
    if (els.size() >= kMaxArgs) {
        er = E_NOMEM;
        return GV();
    }
    for (const std::string& e0 : els) {
        std::string e = e0;
        const size_t eq = find_outside(e, 0, '=');
        if (eq != std::string::npos) {
            c.names.push_back(e.substr(0, eq));
            c.has_name.push_back(true);
            e = e.substr(eq + 1);
        } else {
            c.names.emplace_back();
            c.has_name.push_back(false);
        }
        // only the last argument's error reaches the caller
        c.args.push_back(eval(e));
    }
    if (er != OK) return GV();
    const size_t n = c.args.size();
    auto num_arg = [&](double (*f)(double)) -> GV {
        if (n < 1) {
            er = E_LACK_TOKENS;
            return GV();
        }
        if (c.args[0].t != GV::Num) {
            er = E_BAD_TYPE;
            return GV();
        }
        return GV::num(f(c.args[0].x));
    };
    if (c.name == "vec") {
        if (n < 3) {
            er = E_LACK_TOKENS;
            return GV();
        }
        if (c.args[0].t != GV::Num || c.args[1].t != GV::Num || c.args[2].t != GV::Num) {
            er = E_BAD_TOKEN;
            return GV();
        }
        return GV::vec({c.args[0].x, c.args[1].x, c.args[2].x});
    }
    if (c.name == "range") {
        // range(n), range(a, b) (integers), range(a, b, step)
        GV r;
        r.t = GV::List;
        if (n == 0) {
            er = E_LACK_TOKENS;
            return GV();
        }
        for (size_t q = 0; q < std::min<size_t>(n, 3); ++q)
            if (c.args[q].t != GV::Num) {
                er = E_BAD_TYPE;
                return GV();
            }
        if (n == 1) {
            if (c.args[0].x < 0) {
                er = E_BAD_VALUE;
                return GV();
            }
            for (size_t q = 0; q < (size_t)c.args[0].x; ++q) r.l.push_back(GV::num((double)q));
        } else if (n == 2) {
            const int a = (int)c.args[0].x, bnd = (int)c.args[1].x;
            if (bnd < a) {
                er = E_BAD_VALUE;
                return GV();
            }
            for (int q = a; q < bnd; ++q) r.l.push_back(GV::num(q));
        } else {
            const double a = c.args[0].x, bnd = c.args[1].x, inc = c.args[2].x;
            if (bnd < a || inc == 0) {
                er = E_BAD_VALUE;
                return GV();
            }
            const size_t cnt = (size_t)((bnd - a) / inc);
            for (size_t q = 0; q < cnt; ++q) r.l.push_back(GV::num(q * inc + a));
        }
        return r;
    }
    if (c.name == "linspace") {
        if (n < 3) {
            er = E_LACK_TOKENS;
            return GV();
        }
        if (c.args[0].t != GV::Num || c.args[1].t != GV::Num || c.args[2].t != GV::Num) {
            er = E_BAD_TYPE;
            return GV();
        }
        const size_t cnt = (size_t)c.args[2].x;
        if (cnt < 2) {
            er = E_BAD_VALUE;
            return GV();
        }
        GV r;
        r.t = GV::List;
        const double step = (c.args[1].x - c.args[0].x) / (cnt - 1);
        for (size_t q = 0; q < cnt; ++q) r.l.push_back(GV::num(step * q + c.args[0].x));
        return r;
    }
    if (c.name == "flatten") {
        if (n < 1) {
            er = E_LACK_TOKENS;
            return GV();
        }
        if (c.args[0].t != GV::List) {
            er = E_BAD_VALUE;
            return GV();
        }
        GV r;
        r.t = GV::List;
        std::function<void(const GV&)> walk = [&](const GV& g) {
            for (const GV& e : g.l) {
                if (e.t == GV::List) walk(e);
                else r.l.push_back(e);
            }
        };
        walk(c.args[0]);
        return r;
    }
    if (c.name == "print") {
        for (const GV& a : c.args) {
            if (a.t == GV::Num) std::printf("%f", a.x);
            else if (a.t == GV::Str) std::printf("%s", a.s.c_str());
            else if (a.t == GV::Vec) std::printf("(%f,%f,%f)", a.v.x, a.v.y, a.v.z);
            else std::printf("<object>");
        }
        std::printf("\n");
        std::fflush(stdout);
        return GV();
    }
    if (c.name == "sin") return num_arg(std::sin);
    if (c.name == "cos") return num_arg(std::cos);
    if (c.name == "tan") return num_arg(std::tan);
    if (c.name == "exp") return num_arg(std::exp);
    if (c.name == "sqrt") return num_arg(std::sqrt);
    const GV f = lookup(c.name);
    if (f.t != GV::Func) {
        std::fprintf(stderr, "geom: unrecognized function name %s\n", c.name.c_str());
        er = E_BAD_TYPE;
        return GV();
    }
    if ((size_t)f.min_args > n) return GV();  // too few arguments: an undefined value
    switch (f.fn) {
        case F_UM_TO_L:
        case F_FS_TO_T: {
            if (c.args[0].t != GV::Num) {
                er = E_BAD_TOKEN;
                return GV();
            }
            const GV lp = lookup("l_per_um");
            if (lp.t != GV::Num) {
                er = E_NOT_DEFINED;
                return GV();
            }
            return GV::num(f.fn == F_UM_TO_L ? lp.x * c.args[0].x : 0.299792458 * lp.x * c.args[0].x);
        }
        default:
            return construct(c);
    }
}

const char* const kComp[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};

// The scene constructors (src/cgs_data.cpp:22-407): each returns an instance holding its
// arguments under the names the scene builder reads back.
GV Interp::construct(const Call& c) {
    const size_t n = c.args.size();
    const int fn = lookup(c.name).fn;
    const auto& a = c.args;
    auto need_num = [&](size_t i) { return a[i].t == GV::Num; };
    switch (fn) {
        case F_GAUSSIAN:
        case F_CW: {
            const bool g = fn == F_GAUSSIAN;
            if (n < (g ? 5u : 4u)) {
                er = E_LACK_TOKENS;
                return GV();
            }
            if (a[0].t != GV::Str) {
                er = E_BAD_TOKEN;
                return GV();
            }
            for (size_t i = 1; i < (g ? 5u : 4u); ++i)
                if (!need_num(i)) {
                    er = E_BAD_TOKEN;
                    return GV();
                }
            GV r = GV::inst(g ? "Gaussian_source" : "CW_source");
            int comp = 0;
            for (int q = 0; q < 6; ++q)
                if (a[0].s == kComp[q]) comp = q;
            r.put("component", GV::num(comp));
            r.put("wavelength", a[1]);
            r.put("amplitude", a[2]);
            if (g) {
                r.put("width", a[3]);
                r.put("phase", a[4]);
                r.put("cutoff", GV::num(5));
                r.put("start_time", GV::num(5));
                for (size_t i = 5; i < n; ++i) {
                    if (!c.has_name[i]) continue;
                    if (c.names[i] == "cutoff") r.set("cutoff", a[i]);
                    else if (c.names[i] == "start_time") r.set("start_time", a[i]);
                }
            } else {
                r.put("start_time", a[3]);
                r.put("slowness", GV::num(5));
                r.put("end_time", (n > 4 && a[4].t == GV::Num) ? a[4] : GV::num(std::numeric_limits<double>::max()));
                for (size_t i = 5; i < n; ++i)
                    if (c.has_name[i] && c.names[i] == "slowness") r.set("slowness", a[i]);
            }
            r.put("region", a[n - 1]);
            return r;
        }
        case F_BOX: {
            Vec3 p1, p2;
            if ((er = to_vec(a[0], &p1)) != OK || (er = to_vec(a[1], &p2)) != OK) return GV();
            GV r = GV::inst("Box");
            r.put("pt_1", GV::vec(p1));
            r.put("pt_2", GV::vec(p2));
            return r;
        }
        case F_PLANE: {
            Vec3 nrm;
            double off = 0;
            if (n < 3) {
                if ((er = to_vec(a[0], &nrm)) != OK || a[1].t != GV::Num) {
                    er = E_BAD_VALUE;
                    return GV();
                }
                off = a[1].x;
            } else {
                Vec3 p[3];
                for (int q = 0; q < 3; ++q)
                    if ((er = to_vec(a[q], &p[q])) != OK) return GV();
                // normal of the plane through three points, and its offset
                const Shape sh = Plane(p[0], p[1], p[2]);
                nrm = sh.a;
                off = sh.offset;
            }
            GV r = GV::inst("Plane");
            r.put("normal", GV::vec(nrm));
            r.put("offset", GV::num(off));
            return r;
        }
        case F_SPHERE: {
            Vec3 ce;
            if ((er = to_vec(a[0], &ce)) != OK) return GV();
            if (!need_num(1)) {
                er = E_BAD_VALUE;
                return GV();
            }
            GV r = GV::inst("Sphere");
            r.put("center", GV::vec(ce));
            r.put("radius", a[1]);
            return r;
        }
        case F_CYLINDER: {
            if (n < 3) {
                er = E_LACK_TOKENS;
                return GV();
            }
            Vec3 ce;
            if ((er = to_vec(a[0], &ce)) != OK) return GV();
            if (!need_num(1) || !need_num(2)) {
                er = E_BAD_VALUE;
                return GV();
            }
            GV r = GV::inst("Cylinder");
            r.put("center", GV::vec(ce));
            r.put("h", a[1]);
            r.put("r1", a[2]);
            r.put("r2", (n > 3 && a[3].t == GV::Num) ? a[3] : a[2]);
            return r;
        }
        case F_COMPOSITE: {
            int e = OK;
            const GV vl = to_list(a[n - 1], &e);
            if (e != OK || vl.t != GV::List) {
                er = E_BAD_VALUE;
                return GV();
            }
            GV r = GV::inst("Composite");
            for (size_t i = 0; i + 1 < n; ++i)
                if (c.has_name[i]) r.put(c.names[i], a[i]);
            r.put("geometry", vl);
            return r;
        }
        case F_UNION:
        case F_INTERSECT:
        case F_COMPLEMENT: {
            int e = OK;
            const GV vl = to_list(a[0], &e);
            if (e != OK || vl.t != GV::List) {
                er = E_BAD_VALUE;
                return GV();
            }
            GV r = GV::inst(fn == F_UNION ? "Union" : fn == F_INTERSECT ? "Intersect" : "Complement");
            r.put("geometry", vl);
            return r;
        }
        case F_DIFFERENCE: {
            // Intersect([first, Complement(rest)])
            int e = OK;
            const GV vl = to_list(a[0], &e);
            if (e != OK || vl.t != GV::List || vl.l.size() < 2) {
                er = E_BAD_VALUE;
                return GV();
            }
            GV rest;
            rest.t = GV::List;
            rest.l.assign(vl.l.begin() + 1, vl.l.end());
            GV comp = GV::inst("Complement");
            comp.put("geometry", rest);
            GV both;
            both.t = GV::List;
            both.l = {vl.l[0], comp};
            GV r = GV::inst("Intersect");
            r.put("geometry", both);
            return r;
        }
        case F_ROTATE: {
            if (n < 3) {
                er = E_LACK_TOKENS;
                return GV();
            }
            if (!need_num(0)) {
                er = E_BAD_TYPE;
                return GV();
            }
            Vec3 ax;
            if ((er = to_vec(a[1], &ax)) != OK) return GV();
            int e = OK;
            const GV vg = to_list(a[2], &e);
            if (e != OK) {
                er = e;
                return GV();
            }
            GV r = GV::inst("Rotate");
            r.put("axis", GV::vec(ax));
            r.put("theta", a[0]);
            r.put("geometry", vg);
            return r;
        }
        case F_SNAPSHOT: {
            if (a[0].t != GV::Str) {
                er = E_BAD_VALUE;
                return GV();
            }
            Vec3 cam;
            if ((er = to_vec(a[1], &cam)) != OK) return GV();
            GV r = GV::inst("snapshot");
            r.put("fname", a[0]);
            r.put("cam_v", GV::vec(cam));
            return r;
        }
        case F_MONITORS: {
            if (a[0].t != GV::List) {
                er = E_BAD_TYPE;
                return GV();
            }
            GV r = GV::inst("monitor");
            r.put("locations", a[0]);
            return r;
        }
        default:
            return GV();
    }
}

// One statement starting at the beginning of line *line (read_single_line, src/cgs_read.cpp:2114-
// 2210): "name = expression" or a bare expression; '//' starts a comment; a bracket or quote
// runs on to its partner, across lines if need be, in which case the statement ends there.
int Interp::read_statement(const std::vector<std::string>& lines, size_t* line, size_t* off) {
    const std::string& L = lines[*line];
    const size_t len = L.size();
    std::string buf;
    bool started = false, has_lval = false;
    std::string lval;
    size_t rval = 0;
    size_t pos = 0;
    for (pos = 0;; ++pos) {
        if (pos >= len || L[pos] == 0) break;
        started = true;
        const char c = L[pos];
        if (c == '/' && pos > 0 && L[pos - 1] == '/') {
            if (buf.size() == 1) started = false;
            else pos = len;
            buf.pop_back();
            break;
        }
        if (c == '*' && pos > 0 && L[pos - 1] == '/') {
            std::fprintf(stderr, "geom: block comments are not supported\n");
            return E_BAD_SYNTAX;
        }
        if (c == '=') {
            lval = trim(buf.substr(0, buf.find('\0')));
            has_lval = true;
            buf += '\0';
            rval = buf.size();
            continue;
        }
        if (c == '(' || c == '[' || c == '{' || c == '"' || c == '\'') {
            const char close = c == '"' || c == '\'' ? c : partner(c);
            // the enclosed text from here to the partner, lines joined with a space after each
            std::string text;
            size_t eline = *line, eoff = pos;
            int depth = 0;
            bool found = false;
            for (size_t ln = *line; ln < lines.size() && !found; ++ln) {
                const std::string& S = lines[ln];
                size_t j = ln == *line ? pos : 0;
                const size_t j0 = j;
                for (; j < S.size(); ++j) {
                    const char ch = S[j];
                    if (close == c) {
                        if (ch == c && (j == 0 || S[j - 1] != '\\')) {
                            depth = 1 - depth;
                            if (depth == 0) {
                                found = true;
                                ++j;
                                break;
                            }
                        }
                    } else if (ch == c) {
                        ++depth;
                    } else if (ch == close) {
                        if (--depth <= 0) {
                            found = true;
                            ++j;
                            break;
                        }
                    }
                }
                text += S.substr(j0, j - j0);
                if (!text.empty()) text += ' ';
                eline = found ? ln : ln + 1;
                eoff = j;
            }
            if (!found) eline = lines.size();
            if (eline == *line && eoff < len) {
                buf += text.substr(0, eoff - pos);
                pos = eoff;
                // the character right after the block is taken as it is
            } else {
                buf += text;
                *line = eline;
                *off = eoff;
                goto done;
            }
        }
        buf += L[pos];
    }
    *off = pos;
done:
    if (!started) {
        *line += 1;
        *off = 0;
        return OK;
    }
    const GV v = eval(buf.substr(rval));
    if (er != OK) return er;
    emplace(has_lval ? &lval : nullptr, v);
    return OK;
}

// read_from_lines (src/cgs_read.cpp:2219-2256): statements in order; a statement that ends
// inside a line is followed by a fresh read of that line from its start.
int Interp::read_lines(const std::vector<std::string>& lines) {
    size_t line = 0, off = 0;
    while (line < lines.size()) {
        if (find_word(lines[line], 0, "def") != std::string::npos) {
            std::fprintf(stderr, "geom: function definitions are not supported\n");
            return E_BAD_SYNTAX;
        }
        const int e = read_statement(lines, &line, &off);
        if (e != OK) return e;
        if (line >= lines.size()) break;
        if (off >= lines[line].size()) {
            if (line + 1 >= lines.size()) break;
            ++line;
            off = 0;
        }
    }
    return OK;
}

// lines of a file: split at newlines and ';', empty lines dropped (src/cgs_read.cpp:343-402)
std::vector<std::string> file_lines(const std::string& text) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : text) {
        if (c == '\n' || c == ';') {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

// geometry list entries -> shapes for assemble(); an entry that is no instance is kept as an
// empty placeholder (skipped, but it counts for which entry is last), an instance the tree cannot
// use as a "Stop" (the tree stops reading the list there) (src/cgs.cpp:122-270)
Shape shape_of(const GV& g) {
    Shape s;
    if (g.t != GV::Inst) return s;
    const std::string ty = g.type_name();
    Vec3 v1, v2;
    if (ty == "Box" && to_vec(g.get("pt_1"), &v1) == OK && to_vec(g.get("pt_2"), &v2) == OK) return Box(v1, v2);
    if (ty == "Plane" && g.get("offset").t == GV::Num && to_vec(g.get("normal"), &v1) == OK)
        return Plane(v1, g.get("offset").x);
    if (ty == "Sphere" && g.get("radius").t == GV::Num && to_vec(g.get("center"), &v1) == OK)
        return Sphere(v1, g.get("radius").x);
    if (ty == "Cylinder" && g.get("h").t == GV::Num && g.get("r1").t == GV::Num && g.get("r2").t == GV::Num &&
        to_vec(g.get("center"), &v1) == OK)
        return Cylinder(v1, g.get("h").x, g.get("r1").x, g.get("r2").x);
    if (ty == "Union" || ty == "Intersect" || ty == "Complement" || ty == "Rotate") {
        const GV gl = g.get("geometry");
        std::vector<Shape> list;
        if (gl.t == GV::List)
            for (const GV& e : gl.l) list.push_back(shape_of(e));
        if (ty == "Union" && gl.t == GV::List) return Union(list);
        if (ty == "Intersect" && gl.t == GV::List) return Intersect(list);
        if (ty == "Complement") return Complement(list);
        if (ty == "Rotate") {
            const GV th = g.get("theta"), ax = g.get("axis");
            if (th.t == GV::Num && ax.t == GV::Vec) return Rotate(th.x, ax.v, list);
            return s;  // a rotation without angle or axis adds nothing
        }
    }
    s.type = "Stop";
    return s;
}

Value value_of(const GV& g) {
    Value v;
    if (g.t == GV::Num) {
        v.type = Value::Num;
        v.x = g.x;
    } else if (g.t == GV::Str) {
        v.type = Value::Str;
        v.s = g.s;
    } else if (g.t == GV::List) {
        v.type = Value::List;
        for (const GV& e : g.l) v.l.push_back(value_of(e));
    } else if (g.t == GV::Vec) {
        v.type = Value::List;
        for (double x : {g.v.x, g.v.y, g.v.z}) {
            Value e;
            e.type = Value::Num;
            e.x = x;
            v.l.push_back(e);
        }
    }
    return v;
}

}  // namespace

Scene Scene::from_geom_file(const std::string& path, const Settings& s, int* parse_error) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open geometry file " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return from_geom_text(ss.str(), s, parse_error);
}

Scene Scene::from_geom_text(const std::string& text, const Settings& s, int* parse_error) {
    Interp in;
    // context_from_settings (src/disp.cpp:28-50) and setup_geometry_context
    // (src/cgs_data.cpp:409-453)
    in.define("pi", GV::num(M_PI));
    in.define("pml_thickness", GV::num(s.pml_thickness));
    in.define("sim_length", GV::num(s.len));
    in.define("length", GV::num(2 * s.pml_thickness + s.len));
    in.define("l_per_um", GV::num(s.um_scale));
    in.define("out_dir", GV::str(s.out_dir));
    in.define_func("um_to_l", F_UM_TO_L, 1);
    in.define_func("fs_to_t", F_FS_TO_T, 1);
    int er = OK;
    if (!s.user_opts.empty()) {
        std::vector<std::string> opts;
        std::string cur;
        for (char c : s.user_opts) {
            if (c == ';') {
                opts.push_back(cur);
                cur.clear();
            } else {
                cur += c;
            }
        }
        opts.push_back(cur);
        // the reference ignores the status of the option statements
        (void)in.read_lines(opts);
    }
    in.define_func("Gaussian_source", F_GAUSSIAN, 6);
    in.define_func("CW_source", F_CW, 5);
    in.define_func("Box", F_BOX, 2);
    in.define_func("Plane", F_PLANE, 2);
    in.define_func("Sphere", F_SPHERE, 2);
    in.define_func("Cylinder", F_CYLINDER, 2);
    in.define_func("Composite", F_COMPOSITE, 1);
    in.define_func("Union", F_UNION, 1);
    in.define_func("Intersect", F_INTERSECT, 1);
    in.define_func("Complement", F_COMPLEMENT, 1);
    in.define_func("Rotate", F_ROTATE, 1);
    in.define_func("Difference", F_DIFFERENCE, 1);
    in.define_func("snapshot", F_SNAPSHOT, 2);
    in.define_func("monitors", F_MONITORS, 1);
    const size_t first = in.st.size();
    er = in.read_lines(file_lines(text));

    Scene sc;
    // Composite instances become roots, newest first (scene::read_file, src/cgs.cpp:510-570)
    if (er == OK) {
        for (size_t i = in.st.size(); i > first; --i) {
            const GV& g = in.st[i - 1].v;
            if (g.t != GV::Inst || g.type_name() != "Composite") continue;
            Combine cmb = Combine::Union;
            const GV ct = g.get("combine_type");
            if (ct.t == GV::Str && (ct.s == "intersect" || ct.s == "Intersect")) cmb = Combine::Intersect;
            Root root;
            // metadata: the named arguments (between the type and the geometry); the first of
            // a repeated name wins
            for (size_t q = 1; q + 1 < g.f->size(); ++q) {
                const auto& kv = (*g.f)[q];
                if (!root.metadata.count(kv.first)) root.metadata[kv.first] = value_of(kv.second);
            }
            std::vector<Shape> list;
            const GV geo = g.f->back().second;
            if (geo.t == GV::List)
                for (const GV& e : geo.l) list.push_back(shape_of(e));
            root.tree = assemble(cmb, list);
            sc.roots.push_back(std::move(root));
        }
    }
    // sources and monitor clusters in statement order; every numeric entry of the context
    static const char* keys[9] = {"component", "wavelength", "amplitude", "width", "phase",
                                  "cutoff", "start_time", "end_time", "slowness"};
    for (const Entry& e : in.st) {
        const GV& g = e.v;
        if (g.t == GV::Num) sc.context_numbers.emplace_back(e.named ? e.name : std::string(), g.x);
        if (g.t != GV::Inst) continue;
        const std::string ty = g.type_name();
        if (ty == "Gaussian_source" || ty == "CW_source") {
            SourceSpec sp;
            sp.gaussian = ty == "Gaussian_source";
            double* dst[9] = {&sp.component, &sp.wavelength, &sp.amplitude, &sp.width, &sp.phase,
                              &sp.cutoff, &sp.start_time, &sp.end_time, &sp.slowness};
            for (int q = 0; q < 9; ++q) {
                const GV v = g.get(keys[q]);
                if (v.t == GV::Num) {
                    sp.has[q] = true;
                    *dst[q] = v.x;
                }
            }
            const GV reg = g.get("region");
            if (reg.t == GV::Inst && reg.type_name() == "Box" && to_vec(reg.get("pt_1"), &sp.box[0]) == OK &&
                to_vec(reg.get("pt_2"), &sp.box[1]) == OK)
                sp.has_box = true;
            sc.sources.push_back(sp);
        } else if (ty == "monitor") {
            std::vector<Vec3> pts;
            const GV locs = g.get("locations");
            if (locs.t == GV::List)
                for (const GV& p : locs.l) {
                    Vec3 v;
                    if (to_vec(p, &v) == OK) pts.push_back(v);
                }
            sc.monitor_clusters.push_back(std::move(pts));
        }
    }
    if (parse_error) *parse_error = er;
    return sc;
}

Scene load_scene(const std::string& path, const Settings& s) {
    if (path.size() >= 5 && path.compare(path.size() - 5, 5, ".json") == 0) return Scene::from_json_file(path);
    int er = 0;
    Scene sc = Scene::from_geom_file(path, s, &er);
    if (er != 0) throw std::runtime_error("Scene parsing failed (parse error " + std::to_string(er) + " in " + path + ")");
    return sc;
}

}  // namespace sjfh
