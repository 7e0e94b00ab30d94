// io.cpp - readers of the reference's input files declared in include/llg_io.h (no GPU code).
//
// The configuration and exchange files are libconfig documents: settings "name = value;"
// or "name : value;", groups { }, arrays [ ] of scalars, lists ( ) of any values, strings
// with escapes (adjacent strings concatenate), integers (decimal or 0x hex, optional L
// suffix), floats, booleans (true/false in any case) and #, // and /* */ comments.  The
// parser below covers that grammar, builds a tree, and answers dot-path lookups.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <unordered_map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/llg_capi.h"
#include "../../include/llg_io.h"

namespace sllg {
void set_error(const char* fmt, ...);  // capi.hip
}

#define IO_FAIL(code, ...)                 \
    do {                                   \
        ::sllg::set_error(__VA_ARGS__);    \
        return (code);                     \
    } while (0)

struct CfgNode {
    int type = LLG_CFG_MISSING;
    std::string name;
    double num = 0.0;
    bool b = false;
    std::string str;
    std::vector<CfgNode> kids;
};

struct llg_cfg {
    CfgNode root;
};

namespace {

class Parser {
public:
    explicit Parser(const std::string& t) : s_(t) {}

    bool parse(CfgNode& root, std::string& err) {
        root.type = LLG_CFG_GROUP;
        try {
            skip();
            while (p_ < s_.size()) {
                root.kids.emplace_back();
                setting(root.kids.back());
                skip();
            }
        } catch (const std::string& e) {
            err = e;
            return false;
        }
        return true;
    }

private:
    const std::string& s_;
    size_t p_ = 0;
    int line_ = 1;

    [[noreturn]] void fail(const char* what) {
        std::ostringstream o;
        o << "config syntax error at line " << line_ << ": " << what;
        throw o.str();
    }
    int peek() const { return p_ < s_.size() ? (unsigned char)s_[p_] : -1; }
    int get() {
        if (p_ >= s_.size()) fail("unexpected end of input");
        const int c = (unsigned char)s_[p_++];
        if (c == '\n') ++line_;
        return c;
    }
    void skip() {
        for (;;) {
            while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) get();
            if (p_ >= s_.size()) return;
            if (s_[p_] == '#' || (s_[p_] == '/' && p_ + 1 < s_.size() && s_[p_ + 1] == '/')) {
                while (p_ < s_.size() && s_[p_] != '\n') ++p_;
            } else if (s_[p_] == '/' && p_ + 1 < s_.size() && s_[p_ + 1] == '*') {
                p_ += 2;
                while (p_ + 1 < s_.size() && !(s_[p_] == '*' && s_[p_ + 1] == '/')) get();
                if (p_ + 1 >= s_.size()) fail("unterminated comment");
                p_ += 2;
            } else {
                return;
            }
        }
    }
    static bool name_start(int c) { return std::isalpha(c) || c == '*'; }
    static bool name_char(int c) { return std::isalnum(c) || c == '_' || c == '-' || c == '*'; }

    void setting(CfgNode& n) {
        if (!name_start(peek())) fail("setting name expected");
        while (p_ < s_.size() && name_char(peek())) n.name.push_back((char)get());
        skip();
        const int c = peek();
        if (c != '=' && c != ':') fail("'=' or ':' expected after a setting name");
        get();
        skip();
        value(n);
        skip();
        if (peek() == ';' || peek() == ',') get();
    }

    void value(CfgNode& n) {
        const int c = peek();
        if (c == '{') {
            get();
            n.type = LLG_CFG_GROUP;
            skip();
            while (peek() != '}') {
                if (peek() < 0) fail("unterminated group");
                n.kids.emplace_back();
                setting(n.kids.back());
                skip();
            }
            get();
        } else if (c == '[' || c == '(') {
            const int close = c == '[' ? ']' : ')';
            get();
            n.type = c == '[' ? LLG_CFG_ARRAY : LLG_CFG_LIST;
            skip();
            while (peek() != close) {
                if (peek() < 0) fail("unterminated array or list");
                n.kids.emplace_back();
                value(n.kids.back());
                if (n.type == LLG_CFG_ARRAY && (n.kids.back().type < LLG_CFG_INT))
                    fail("arrays hold scalars only");
                skip();
                if (peek() == ',') {
                    get();
                    skip();
                } else if (peek() != close) {
                    fail("',' expected between elements");
                }
            }
            get();
        } else if (c == '"') {
            n.type = LLG_CFG_STRING;
            for (;;) {
                get();
                while (peek() != '"') {
                    int ch = get();
                    if (ch == '\\') {
                        const int e = get();
                        ch = e == 'n' ? '\n' : e == 't' ? '\t' : e == 'r' ? '\r' : e == 'f' ? '\f' : e;
                    }
                    n.str.push_back((char)ch);
                }
                get();
                skip();
                if (peek() != '"') break;  // adjacent strings concatenate
            }
        } else if (std::isdigit(c) || c == '+' || c == '-' || c == '.') {
            number(n);
        } else if (std::isalpha(c)) {
            std::string w;
            while (p_ < s_.size() && std::isalpha(peek())) w.push_back((char)std::tolower(get()));
            if (w == "true" || w == "false") {
                n.type = LLG_CFG_BOOL;
                n.b = w == "true";
                n.num = n.b ? 1.0 : 0.0;
            } else {
                fail("value expected");
            }
        } else {
            fail("value expected");
        }
    }

    void number(CfgNode& n) {
        const size_t b = p_;
        while (p_ < s_.size() && (std::isalnum(peek()) || peek() == '.' || peek() == '+' || peek() == '-')) {
            // a sign is part of the number only at its start or after an exponent marker
            if ((peek() == '+' || peek() == '-') && p_ > b && std::tolower(s_[p_ - 1]) != 'e') break;
            ++p_;
        }
        std::string t = s_.substr(b, p_ - b);
        while (!t.empty() && (t.back() == 'L' || t.back() == 'l')) t.pop_back();  // 64-bit suffix
        const bool hex = t.size() > 2 && (t[0] == '0' || ((t[0] == '-' || t[0] == '+') && t[1] == '0')) &&
                         (t.find('x') != std::string::npos || t.find('X') != std::string::npos);
        char* end = nullptr;
        errno = 0;
        if (hex) {
            n.type = LLG_CFG_INT;
            n.num = (double)std::strtoll(t.c_str(), &end, 16);
        } else if (t.find_first_of(".eE") != std::string::npos) {
            n.type = LLG_CFG_FLOAT;
            n.num = std::strtod(t.c_str(), &end);
        } else {
            n.type = LLG_CFG_INT;
            n.num = (double)std::strtoll(t.c_str(), &end, 10);
        }
        if (t.empty() || !end || *end != '\0' || errno == ERANGE) fail("malformed number");
    }
};

const CfgNode* lookup(const llg_cfg* cfg, const char* path) {
    if (!cfg || !path) return nullptr;
    const CfgNode* n = &cfg->root;
    std::string p(path);
    if (p.empty()) return n;
    size_t b = 0;
    while (b <= p.size()) {
        const size_t e = p.find('.', b);
        const std::string part = p.substr(b, e == std::string::npos ? std::string::npos : e - b);
        if (part.empty()) return nullptr;
        if (n->type != LLG_CFG_GROUP) return nullptr;
        const CfgNode* next = nullptr;
        for (const CfgNode& k : n->kids)
            if (k.name == part) {
                next = &k;
                break;
            }
        if (!next) return nullptr;
        n = next;
        if (e == std::string::npos) break;
        b = e + 1;
    }
    return n;
}

int number_of(const CfgNode* n, uint32_t index, double* out, const char* path) {
    if (!n) IO_FAIL(LLG_E_INVALID, "setting %s not found", path);
    if (n->type == LLG_CFG_ARRAY || n->type == LLG_CFG_LIST) {
        if (index >= n->kids.size()) IO_FAIL(LLG_E_INVALID, "setting %s has no element %u", path, index);
        n = &n->kids[index];
    } else if (index != 0) {
        IO_FAIL(LLG_E_INVALID, "setting %s is not an array", path);
    }
    if (n->type != LLG_CFG_INT && n->type != LLG_CFG_FLOAT && n->type != LLG_CFG_BOOL)
        IO_FAIL(LLG_E_INVALID, "setting %s is not a number", path);
    *out = n->num;
    return 0;
}

// Buffered whitespace tokenizer for the large CSR matrix file.
class Tokens {
public:
    explicit Tokens(FILE* f) : f_(f), buf_(1 << 20) {}
    bool next(std::string& t) {
        t.clear();
        for (;;) {
            if (pos_ >= len_ && !fill()) return !t.empty();
            const char c = buf_[pos_];
            if (std::isspace((unsigned char)c)) {
                ++pos_;
                if (!t.empty()) return true;
            } else {
                t.push_back(c);
                ++pos_;
            }
        }
    }

private:
    bool fill() {
        len_ = std::fread(buf_.data(), 1, buf_.size(), f_);
        pos_ = 0;
        return len_ > 0;
    }
    FILE* f_;
    std::vector<char> buf_;
    size_t pos_ = 0, len_ = 0;
};

bool parse_u64(const std::string& t, uint64_t* v) {
    if (t.empty() || !std::isdigit((unsigned char)t[0])) return false;
    char* end = nullptr;
    errno = 0;
    *v = std::strtoull(t.c_str(), &end, 10);
    return *end == '\0' && errno != ERANGE;
}
bool parse_f64(const std::string& t, double* v) {
    char* end = nullptr;
    errno = 0;
    *v = std::strtod(t.c_str(), &end);
    return !t.empty() && *end == '\0';
}

struct FileCloser {
    void operator()(FILE* f) const {
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" {

int llg_cfg_parse_string(const char* text, llg_cfg** out) {
    if (!text || !out) IO_FAIL(LLG_E_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<llg_cfg> c(new llg_cfg);
    const std::string s(text);
    Parser p(s);
    std::string err;
    if (!p.parse(c->root, err)) IO_FAIL(LLG_E_INVALID, "%s", err.c_str());
    *out = c.release();
    return 0;
}

int llg_cfg_parse_file(const char* path, llg_cfg** out) {
    if (!path || !out) IO_FAIL(LLG_E_INVALID, "null argument");
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) IO_FAIL(LLG_E_INVALID, "cannot open %s", path);
    std::ostringstream ss;
    ss << f.rdbuf();
    std::unique_ptr<llg_cfg> c(new llg_cfg);
    const std::string s = ss.str();
    Parser p(s);
    std::string err;
    if (!p.parse(c->root, err)) IO_FAIL(LLG_E_INVALID, "%s: %s", path, err.c_str());
    *out = c.release();
    return 0;
}

void llg_cfg_free(llg_cfg* cfg) { delete cfg; }

int llg_cfg_type(const llg_cfg* cfg, const char* path) {
    const CfgNode* n = lookup(cfg, path);
    return n ? n->type : LLG_CFG_MISSING;
}

int llg_cfg_length(const llg_cfg* cfg, const char* path, uint32_t* len) {
    const CfgNode* n = lookup(cfg, path);
    if (!n || !len) IO_FAIL(LLG_E_INVALID, "setting %s not found", path ? path : "(null)");
    *len = (n->type == LLG_CFG_GROUP || n->type == LLG_CFG_ARRAY || n->type == LLG_CFG_LIST) ? (uint32_t)n->kids.size()
                                                                                               : 1u;
    return 0;
}

int llg_cfg_get_number(const llg_cfg* cfg, const char* path, uint32_t index, double* out) {
    if (!out) IO_FAIL(LLG_E_INVALID, "null argument");
    return number_of(lookup(cfg, path), index, out, path ? path : "(null)");
}

int llg_cfg_get_bool(const llg_cfg* cfg, const char* path, int* out) {
    const CfgNode* n = lookup(cfg, path);
    if (!n || !out) IO_FAIL(LLG_E_INVALID, "setting %s not found", path ? path : "(null)");
    if (n->type != LLG_CFG_BOOL) IO_FAIL(LLG_E_INVALID, "setting %s is not a boolean", path);
    *out = n->b ? 1 : 0;
    return 0;
}

int llg_cfg_get_string(const llg_cfg* cfg, const char* path, char* buf, uint32_t cap, uint32_t* len) {
    const CfgNode* n = lookup(cfg, path);
    if (!n) IO_FAIL(LLG_E_INVALID, "setting %s not found", path ? path : "(null)");
    if (n->type != LLG_CFG_STRING) IO_FAIL(LLG_E_INVALID, "setting %s is not a string", path);
    if (len) *len = (uint32_t)n->str.size();
    if (buf) {
        if (cap < n->str.size() + 1) IO_FAIL(LLG_E_INVALID, "buffer too small for %s", path);
        std::memcpy(buf, n->str.c_str(), n->str.size() + 1);
    }
    return 0;
}

int llg_cfg_child_name(const llg_cfg* cfg, const char* path, uint32_t index, char* buf, uint32_t cap, uint32_t* len) {
    const CfgNode* n = lookup(cfg, path);
    if (!n || n->type != LLG_CFG_GROUP) IO_FAIL(LLG_E_INVALID, "group %s not found", path ? path : "(null)");
    if (index >= n->kids.size()) IO_FAIL(LLG_E_INVALID, "group %s has no setting %u", path, index);
    const std::string& s = n->kids[index].name;
    if (len) *len = (uint32_t)s.size();
    if (buf) {
        if (cap < s.size() + 1) IO_FAIL(LLG_E_INVALID, "buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    }
    return 0;
}

int llg_read_ucf(const char* path, int with_mesh, uint32_t* n_species, uint32_t* n_atoms, llg_ucf_atom* atoms,
                 uint32_t cap) {
    if (!path || !n_species || !n_atoms) IO_FAIL(LLG_E_INVALID, "null argument");
    std::ifstream f(path);
    if (!f) IO_FAIL(LLG_E_INVALID, "could not open unit cell file %s", path);
    long long nms = -1, nauc = -1;
    if (!(f >> nms >> nauc) || nms < 1) IO_FAIL(LLG_E_INVALID, "%s: could not read the species and atom counts", path);
    if (nauc < 1) IO_FAIL(LLG_E_INVALID, "%s: there are no atoms in the unit cell", path);
    *n_species = (uint32_t)nms;
    *n_atoms = (uint32_t)nauc;
    if (!atoms) return 0;
    if (cap < (uint64_t)nauc) IO_FAIL(LLG_E_INVALID, "%s: %lld atoms, room for %u", path, nauc, cap);
    for (long long i = 0; i < nauc; ++i) {
        llg_ucf_atom a{};
        long long sl = -1;
        std::string el;
        bool ok = (bool)(f >> sl >> a.xred[0] >> a.xred[1] >> a.xred[2]);
        if (ok && with_mesh) {
            long long mp[3];
            ok = (bool)(f >> mp[0] >> mp[1] >> mp[2]) && mp[0] >= 0 && mp[1] >= 0 && mp[2] >= 0;
            for (int k = 0; k < 3 && ok; ++k) a.mp[k] = (uint32_t)mp[k];
        }
        ok = ok && (f >> a.mu >> a.lambda >> a.gamma_rel >> el >> a.spin0[0] >> a.spin0[1] >> a.spin0[2] >> a.k1u >>
                    a.k1u_dir[0] >> a.k1u_dir[1] >> a.k1u_dir[2]);
        if (!ok || sl < 0) IO_FAIL(LLG_E_INVALID, "%s: could not read the record of atom %lld", path, i);
        if (sl >= nms) IO_FAIL(LLG_E_INVALID, "%s: atom %lld has sublattice %lld of %lld", path, i, sl, nms);
        if (a.mu < 1e-5 || a.mu > 1e5)
            IO_FAIL(LLG_E_INVALID, "%s: magnetic moment %g of atom %lld outside (1e-5, 1e5)", path, a.mu, i);
        if (a.gamma_rel < 1e-5 || a.gamma_rel > 1e5)
            IO_FAIL(LLG_E_INVALID, "%s: gyromagnetic ratio %g of atom %lld outside (1e-5, 1e5)", path, a.gamma_rel, i);
        a.sublattice = (uint32_t)sl;
        std::snprintf(a.element, sizeof a.element, "%s", el.c_str());
        atoms[i] = a;
    }
    return 0;
}

int llg_read_exch_unitcell(const llg_cfg* exch, uint32_t n_atoms, llg_interaction* out, uint32_t cap,
                           uint32_t* nout) {
    if (!exch || !nout || n_atoms == 0) IO_FAIL(LLG_E_INVALID, "null argument");
    uint32_t cnt = 0;
    char key[64], sub[96];
    for (uint32_t a = 0; a < n_atoms; ++a)
        for (uint32_t b = 0; b < n_atoms; ++b) {
            std::snprintf(key, sizeof key, "exchange_%u_%u", a, b);
            if (llg_cfg_type(exch, key) == LLG_CFG_MISSING) continue;
            std::snprintf(sub, sizeof sub, "%s.Num_Interactions", key);
            double nint = 0;
            if (llg_cfg_get_number(exch, sub, 0, &nint) || nint < 0)
                IO_FAIL(LLG_E_INVALID, "could not read the number of interactions between species %u and %u", a, b);
            // energy unit (the reference defaults to joules with a warning)
            double f1 = 1.0, f2 = 1.0;
            std::snprintf(sub, sizeof sub, "%s.units", key);
            char units[32] = "J";
            if (llg_cfg_type(exch, sub) == LLG_CFG_STRING &&
                llg_cfg_get_string(exch, sub, units, sizeof units, nullptr))
                IO_FAIL(LLG_E_INVALID, "units of %s too long", key);
            const std::string u(units);
            if (u == "mRy") {
                f1 = 2.179872172e-18;
                f2 = 1.0e-3;
            } else if (u == "eV") {
                f1 = 1.602176565e-19;
            } else if (u == "meV") {
                f1 = 1.602176565e-22;
            } else if (u == "Ry") {
                f1 = 2.179872172e-18;
            } else if (!(u == "J" || u == "Joules" || u == "joules")) {
                IO_FAIL(LLG_E_INVALID, "units of exchange energy not recognised: %s", units);
            }
            for (uint32_t k = 0; k < (uint32_t)nint; ++k) {
                llg_interaction it{};
                it.atom_i = a;
                it.atom_j = b;
                it.shell = (int32_t)k;
                for (int j = 0; j < 3; ++j) {
                    std::snprintf(sub, sizeof sub, "%s.UnitCell%u", key, k + 1);
                    double v = 0;
                    if (llg_cfg_get_number(exch, sub, (uint32_t)j, &v))
                        IO_FAIL(LLG_E_INVALID, "could not read vector for interaction %u between %u and %u", k + 1, a,
                                b);
                    it.dc[j] = v < 0 ? (int32_t)(v - 0.5) : (int32_t)(v + 0.5);
                }
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        std::snprintf(sub, sizeof sub, "%s.J%u_%d", key, k + 1, r + 1);
                        double v = 0;
                        if (llg_cfg_get_number(exch, sub, (uint32_t)c, &v))
                            IO_FAIL(LLG_E_INVALID, "could not read exchange interaction %u tensor row %d between %u and %u",
                                    k + 1, r + 1, a, b);
                        v *= f1;
                        v *= f2;
                        it.J[r * 3 + c] = v;
                    }
                if (out) {
                    if (cnt >= cap) IO_FAIL(LLG_E_INVALID, "more than %u interactions", cap);
                    out[cnt] = it;
                }
                ++cnt;
            }
        }
    *nout = cnt;
    return 0;
}

int llg_csr_read(const char* path, uint32_t n, uint32_t* xadj, uint32_t* adjncy, double* vals3, uint64_t* nnz) {
    if (!path || !nnz) IO_FAIL(LLG_E_INVALID, "null argument");
    std::unique_ptr<FILE, FileCloser> f(std::fopen(path, "rb"));
    if (!f) IO_FAIL(LLG_E_INVALID, "could not open CSR exchange matrix %s", path);
    Tokens tk(f.get());
    std::string t;
    uint64_t lns = 0, adjsize = 0;
    if (!tk.next(t) || !parse_u64(t, &lns) || !tk.next(t) || !parse_u64(t, &adjsize))
        IO_FAIL(LLG_E_INVALID, "%s: could not read the header", path);
    if (lns != n) IO_FAIL(LLG_E_INVALID, "%s: %llu spins in the file, %u in the system", path, (unsigned long long)lns, n);
    if (adjsize < 1) IO_FAIL(LLG_E_INVALID, "%s: the neighbour list is empty", path);
    *nnz = adjsize;
    if (!adjncy) return 0;
    if (!xadj || !vals3) IO_FAIL(LLG_E_INVALID, "null argument");
    for (uint32_t i = 0; i <= n; ++i) {
        uint64_t v;
        if (!tk.next(t) || !parse_u64(t, &v) || v > adjsize)
            IO_FAIL(LLG_E_INVALID, "%s: bad row pointer %u", path, i);
        xadj[i] = (uint32_t)v;
        if (i > 0 && xadj[i] < xadj[i - 1]) IO_FAIL(LLG_E_INVALID, "%s: row pointers decrease at row %u", path, i);
    }
    if (xadj[0] != 0 || xadj[n] != adjsize)
        IO_FAIL(LLG_E_INVALID, "%s: row pointers do not span the %llu entries", path, (unsigned long long)adjsize);
    for (uint64_t q = 0; q < adjsize; ++q) {
        uint64_t j;
        double v[3];
        if (!tk.next(t) || !parse_u64(t, &j) || j >= n)
            IO_FAIL(LLG_E_INVALID, "%s: bad column at entry %llu", path, (unsigned long long)q);
        for (int c = 0; c < 3; ++c)
            if (!tk.next(t) || !parse_f64(t, &v[c]))
                IO_FAIL(LLG_E_INVALID, "%s: bad value at entry %llu", path, (unsigned long long)q);
        adjncy[q] = (uint32_t)j;
        for (int c = 0; c < 3; ++c) vals3[(size_t)c * adjsize + q] = v[c];
    }
    return 0;
}

int llg_csr_write(const char* path, uint32_t n, const uint32_t* xadj, const uint32_t* adjncy, const double* vals3,
                  uint64_t nnz) {
    if (!path || !xadj || (nnz && (!adjncy || !vals3))) IO_FAIL(LLG_E_INVALID, "null argument");
    if (xadj[0] != 0 || xadj[n] != nnz) IO_FAIL(LLG_E_INVALID, "row pointers do not span the %llu entries",
                                                (unsigned long long)nnz);
    std::unique_ptr<FILE, FileCloser> f(std::fopen(path, "w"));
    if (!f) IO_FAIL(LLG_E_INVALID, "could not open %s for writing", path);
    FILE* o = f.get();
    std::fprintf(o, "%u\n%llu\n", n, (unsigned long long)nnz);
    for (uint32_t i = 0; i < n; ++i) std::fprintf(o, "%u\t", xadj[i]);
    std::fprintf(o, "%u\n", xadj[n]);
    // values with 17 significant digits, so a read returns the same doubles
    for (uint64_t q = 0; q < nnz; ++q)
        std::fprintf(o, "%u\t%.17g\t%.17g\t%.17g\t", adjncy[q], vals3[q], vals3[nnz + q], vals3[2 * nnz + q]);
    std::fprintf(o, "\n");
    if (std::ferror(o)) IO_FAIL(LLG_E_INVALID, "write error on %s", path);
    return 0;
}

// exch::read4spin + exch::setup4SpinCSR (src/exch_read_4_spin.cpp:29-258).
int llg_four_spin_quartets(const llg_cfg* exch, uint32_t nms, const double* mu_species, uint32_t n,
                           const uint32_t* mesh, const uint32_t* sublattice, const uint32_t grid[3],
                           const uint32_t nk[3], const uint32_t pbc[3], int rounding, uint32_t* xadj, uint32_t* jkl,
                           uint64_t cap, uint64_t* nq, double* jq_species) {
    if (!exch || !nq || !mesh || !sublattice || !grid || !nk || !pbc || !mu_species || nms == 0 || (jkl && !xadj))
        IO_FAIL(LLG_E_INVALID, "null argument");
    if (rounding != LLG_FOUR_ROUND_NEAREST && rounding != LLG_FOUR_ROUND_REFERENCE)
        IO_FAIL(LLG_E_INVALID, "unknown rounding %d", rounding);
    double maxq = 0;
    if (llg_cfg_get_number(exch, "exchange.MaxQuartetPermutations", 0, &maxq) || maxq < 0)
        IO_FAIL(LLG_E_INVALID, "could not read the setting exchange.MaxQuartetPermutations (int)");
    double scale[3] = {1.0, 1.0, 1.0};
    if (llg_cfg_type(exch, "exchange.Scale") != LLG_CFG_MISSING)
        for (uint32_t c = 0; c < 3; ++c)
            if (llg_cfg_get_number(exch, "exchange.Scale", c, &scale[c]))
                IO_FAIL(LLG_E_INVALID, "could not read exchange.Scale[%u]", c);
    // per species: quartet count, interface species, mesh offsets [q][m][c]
    std::vector<uint32_t> nperm(nms);
    std::vector<int64_t> iface(nms);
    std::vector<std::vector<int32_t>> off(nms);
    char key[64], sub[128];
    for (uint32_t s = 0; s < nms; ++s) {
        std::snprintf(key, sizeof key, "FourSpinExchange_%u", s);
        if (llg_cfg_type(exch, key) != LLG_CFG_GROUP)
            IO_FAIL(LLG_E_INVALID, "setting %s does not exist", key);
        double v = 0;
        std::snprintf(sub, sizeof sub, "%s.NumPermutations", key);
        if (llg_cfg_get_number(exch, sub, 0, &v) || v < 0)
            IO_FAIL(LLG_E_INVALID, "could not read the number of quartets of species %u", s);
        if (v > maxq) IO_FAIL(LLG_E_INVALID, "species %u: %g quartets, more than MaxQuartetPermutations %g", s, v, maxq);
        nperm[s] = (uint32_t)v;
        std::snprintf(sub, sizeof sub, "%s.JQ", key);
        if (llg_cfg_get_number(exch, sub, 0, &v)) IO_FAIL(LLG_E_INVALID, "could not read JQ of species %u", s);
        if (jq_species) jq_species[s] = v / (mu_species[s] * 9.27e-24);
        std::snprintf(sub, sizeof sub, "%s.InterfaceSpecies", key);
        if (llg_cfg_get_number(exch, sub, 0, &v)) IO_FAIL(LLG_E_INVALID, "could not read InterfaceSpecies of species %u", s);
        iface[s] = (int64_t)v;
        off[s].resize((size_t)nperm[s] * 9);
        for (uint32_t q = 0; q < nperm[s]; ++q)
            for (uint32_t m = 0; m < 3; ++m) {
                std::snprintf(sub, sizeof sub, "%s.Permute%u_%u", key, q + 1, m + 1);
                for (uint32_t c = 0; c < 3; ++c) {
                    double o = 0;
                    if (llg_cfg_get_number(exch, sub, c, &o)) IO_FAIL(LLG_E_INVALID, "setting %s not found", sub);
                    const double x = o * (double)nk[c] * scale[c];
                    int32_t r;
                    if (rounding == LLG_FOUR_ROUND_REFERENCE)
                        r = (int32_t)(x + 0.5);  // truncation toward zero, as the reference
                    else
                        r = x < 0 ? -(int32_t)(-x + 0.5) : (int32_t)(x + 0.5);
                    off[s][((size_t)q * 3 + m) * 3 + c] = r;
                }
            }
    }
    // mesh point -> spin
    const uint64_t gsize = (uint64_t)grid[0] * grid[1] * grid[2];
    if (gsize == 0 || gsize > (uint64_t)1 << 34) IO_FAIL(LLG_E_INVALID, "bad mesh extent");
    std::vector<int64_t> at(gsize, -1);
    for (uint32_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c)
            if (mesh[3 * (size_t)i + c] >= grid[c]) IO_FAIL(LLG_E_INVALID, "spin %u lies outside the mesh", i);
        if (sublattice[i] >= nms) IO_FAIL(LLG_E_INVALID, "spin %u: species %u of %u", i, sublattice[i], nms);
        const uint64_t g = ((uint64_t)mesh[3 * (size_t)i] * grid[1] + mesh[3 * (size_t)i + 1]) * grid[2] +
                           mesh[3 * (size_t)i + 2];
        if (at[g] >= 0) IO_FAIL(LLG_E_INVALID, "spins %lld and %u share a mesh point", (long long)at[g], i);
        at[g] = i;
    }
    uint64_t cnt = 0;
    if (xadj) xadj[0] = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t sl = sublattice[i];
        for (uint32_t q = 0; q < nperm[sl]; ++q) {
            int64_t who[3];
            bool ok = true;
            for (uint32_t m = 0; m < 3 && ok; ++m) {
                int64_t p[3];
                for (int c = 0; c < 3; ++c) {
                    p[c] = (int64_t)mesh[3 * (size_t)i + c] + off[sl][((size_t)q * 3 + m) * 3 + c];
                    const int64_t G = grid[c];
                    if (p[c] < 0 && pbc[c]) p[c] += G;
                    else if (p[c] >= G && pbc[c]) p[c] -= G;
                    ok = ok && p[c] >= 0 && p[c] < G;
                }
                if (!ok) break;
                const int64_t j = at[((uint64_t)p[0] * grid[1] + (uint64_t)p[1]) * grid[2] + (uint64_t)p[2]];
                ok = j >= 0 && ((int64_t)sublattice[j] == (int64_t)sl || (int64_t)sublattice[j] == iface[sl]);
                who[m] = j;
            }
            if (!ok) continue;
            if (jkl) {
                if (cnt >= cap) IO_FAIL(LLG_E_INVALID, "more than %llu quartets", (unsigned long long)cap);
                for (int m = 0; m < 3; ++m) jkl[3 * cnt + (uint64_t)m] = (uint32_t)who[m];
            }
            ++cnt;
        }
        if (xadj) {
            if (cnt > 0xFFFFFFFFull) IO_FAIL(LLG_E_INVALID, "more than 2^32 quartets");
            xadj[i + 1] = (uint32_t)cnt;
        }
    }
    *nq = cnt;
    return 0;
}

// 4spin_csr_exch_mat.dat (src/exch_read_4_spin.cpp:260-381).
int llg_four_csr_read(const char* path, uint32_t n, uint32_t* xadj, uint32_t* jkl, uint64_t* nq) {
    if (!path || !nq) IO_FAIL(LLG_E_INVALID, "null argument");
    std::unique_ptr<FILE, FileCloser> f(std::fopen(path, "rb"));
    if (!f) IO_FAIL(LLG_E_INVALID, "could not open four-spin CSR matrix %s", path);
    Tokens tk(f.get());
    std::string t;
    uint64_t lns = 0, adjsize = 0;
    if (!tk.next(t) || !parse_u64(t, &lns) || !tk.next(t) || !parse_u64(t, &adjsize))
        IO_FAIL(LLG_E_INVALID, "%s: could not read the header", path);
    if (lns != n) IO_FAIL(LLG_E_INVALID, "%s: %llu spins in the file, %u in the system", path, (unsigned long long)lns, n);
    if (adjsize < 1) IO_FAIL(LLG_E_INVALID, "%s: the quartet list is empty", path);
    *nq = adjsize;
    if (!jkl) return 0;
    if (!xadj) IO_FAIL(LLG_E_INVALID, "null argument");
    for (uint32_t i = 0; i <= n; ++i) {
        uint64_t v;
        if (!tk.next(t) || !parse_u64(t, &v) || v > adjsize) IO_FAIL(LLG_E_INVALID, "%s: bad row pointer %u", path, i);
        xadj[i] = (uint32_t)v;
        if (i > 0 && xadj[i] < xadj[i - 1]) IO_FAIL(LLG_E_INVALID, "%s: row pointers decrease at row %u", path, i);
    }
    if (xadj[0] != 0 || xadj[n] != adjsize)
        IO_FAIL(LLG_E_INVALID, "%s: row pointers do not span the %llu quartets", path, (unsigned long long)adjsize);
    for (uint64_t q = 0; q < 3 * adjsize; ++q) {
        uint64_t j;
        if (!tk.next(t) || !parse_u64(t, &j) || j >= n)
            IO_FAIL(LLG_E_INVALID, "%s: bad spin index at quartet %llu", path, (unsigned long long)(q / 3));
        jkl[q] = (uint32_t)j;
    }
    return 0;
}

int llg_four_csr_write(const char* path, uint32_t n, const uint32_t* xadj, const uint32_t* jkl, uint64_t nq) {
    if (!path || !xadj || (nq && !jkl)) IO_FAIL(LLG_E_INVALID, "null argument");
    if (xadj[0] != 0 || xadj[n] != nq)
        IO_FAIL(LLG_E_INVALID, "row pointers do not span the %llu quartets", (unsigned long long)nq);
    std::unique_ptr<FILE, FileCloser> f(std::fopen(path, "w"));
    if (!f) IO_FAIL(LLG_E_INVALID, "could not open %s for writing", path);
    FILE* o = f.get();
    std::fprintf(o, "%u\n%llu\n", n, (unsigned long long)nq);
    for (uint32_t i = 0; i < n; ++i) std::fprintf(o, "%u\t", xadj[i]);
    std::fprintf(o, "%u\n", xadj[n]);
    for (uint64_t q = 0; q < nq; ++q) std::fprintf(o, "%u\t%u\t%u\t", jkl[3 * q], jkl[3 * q + 1], jkl[3 * q + 2]);
    std::fprintf(o, "\n");
    if (std::ferror(o)) IO_FAIL(LLG_E_INVALID, "write error on %s", path);
    return 0;
}

int llg_read_spins_text(const char* path, uint32_t n, double* sx, double* sy, double* sz) {
    if (!path || !sx || !sy || !sz) IO_FAIL(LLG_E_INVALID, "null argument");
    std::unique_ptr<FILE, FileCloser> f(std::fopen(path, "rb"));
    if (!f) IO_FAIL(LLG_E_INVALID, "could not open spin file %s", path);
    Tokens tk(f.get());
    std::string t;
    uint64_t cnt = 0;
    if (!tk.next(t) || !parse_u64(t, &cnt)) IO_FAIL(LLG_E_INVALID, "%s: could not read the number of spins", path);
    if (cnt != n) IO_FAIL(LLG_E_INVALID, "%s: %llu spins in the file, %u in the system", path, (unsigned long long)cnt, n);
    for (uint32_t i = 0; i < n; ++i) {
        double v[3];
        for (int c = 0; c < 3; ++c)
            if (!tk.next(t) || !parse_f64(t, &v[c])) IO_FAIL(LLG_E_INVALID, "%s: bad spin %u", path, i);
        const double m = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (m < 1e-12) IO_FAIL(LLG_E_INVALID, "%s: spin %u has zero length", path, i);
        sx[i] = v[0] / m;
        sy[i] = v[1] / m;
        sz[i] = v[2] / m;
    }
    return 0;
}

int llg_write_spins_text(const char* path, uint32_t n, const double* sx, const double* sy, const double* sz) {
    if (!path || !sx || !sy || !sz) IO_FAIL(LLG_E_INVALID, "null argument");
    std::unique_ptr<FILE, FileCloser> f(std::fopen(path, "w"));
    if (!f) IO_FAIL(LLG_E_INVALID, "could not open %s for writing", path);
    std::fprintf(f.get(), "%u\n", n);
    for (uint32_t i = 0; i < n; ++i) std::fprintf(f.get(), "%.17g\t%.17g\t%.17g\n", sx[i], sy[i], sz[i]);
    if (std::ferror(f.get())) IO_FAIL(LLG_E_INVALID, "write error on %s", path);
    return 0;
}

int llg_read_grid(const char* path, uint32_t n, const uint32_t* mesh, double* sx, double* sy, double* sz) {
    if (!path || !mesh || !sx || !sy || !sz) IO_FAIL(LLG_E_INVALID, "null argument");
    std::unordered_map<uint64_t, uint32_t> site;
    site.reserve(n * 2);
    auto key = [](uint64_t x, uint64_t y, uint64_t z) { return (x << 42) | (y << 21) | z; };
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t* m = mesh + 3 * (size_t)i;
        if (m[0] >= (1u << 21) || m[1] >= (1u << 21) || m[2] >= (1u << 21))
            IO_FAIL(LLG_E_INVALID, "mesh coordinate of spin %u too large", i);
        site[key(m[0], m[1], m[2])] = i;
    }
    std::unique_ptr<FILE, FileCloser> f(std::fopen(path, "rb"));
    if (!f) IO_FAIL(LLG_E_INVALID, "could not open grid file %s", path);
    Tokens tk(f.get());
    std::string t;
    uint64_t cnt = 0;
    if (!tk.next(t) || !parse_u64(t, &cnt)) IO_FAIL(LLG_E_INVALID, "%s: could not read the number of spins", path);
    if (cnt != n) IO_FAIL(LLG_E_INVALID, "%s: %llu spins in the file, %u in the system", path, (unsigned long long)cnt, n);
    std::vector<uint8_t> seen(n, 0);
    for (uint32_t line = 0; line < n; ++line) {
        uint64_t c[3];
        double v[3];
        for (int a = 0; a < 3; ++a)
            if (!tk.next(t) || !parse_u64(t, &c[a])) IO_FAIL(LLG_E_INVALID, "%s: bad grid point on record %u", path, line);
        for (int a = 0; a < 3; ++a)
            if (!tk.next(t) || !parse_f64(t, &v[a])) IO_FAIL(LLG_E_INVALID, "%s: bad spin on record %u", path, line);
        auto it = (c[0] < (1u << 21) && c[1] < (1u << 21) && c[2] < (1u << 21)) ? site.find(key(c[0], c[1], c[2]))
                                                                                : site.end();
        if (it == site.end())
            IO_FAIL(LLG_E_INVALID, "%s: no spin at site (%llu, %llu, %llu)", path, (unsigned long long)c[0],
                    (unsigned long long)c[1], (unsigned long long)c[2]);
        const uint32_t i = it->second;
        const double m = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (m < 1e-12) IO_FAIL(LLG_E_INVALID, "%s: spin of record %u has zero length", path, line);
        sx[i] = v[0] / m;
        sy[i] = v[1] / m;
        sz[i] = v[2] / m;
        seen[i] = 1;
    }
    for (uint32_t i = 0; i < n; ++i)
        if (!seen[i]) IO_FAIL(LLG_E_INVALID, "%s: spin %u is not in the file", path, i);
    return 0;
}

int llg_write_grid(const char* path, uint32_t n, const uint32_t* mesh, const double* sx, const double* sy,
                   const double* sz) {
    if (!path || !mesh || !sx || !sy || !sz) IO_FAIL(LLG_E_INVALID, "null argument");
    std::unique_ptr<FILE, FileCloser> f(std::fopen(path, "w"));
    if (!f) IO_FAIL(LLG_E_INVALID, "could not open %s for writing", path);
    std::fprintf(f.get(), "%u\n", n);
    for (uint32_t i = 0; i < n; ++i)
        std::fprintf(f.get(), "%u\t%u\t%u\t%.17g\t%.17g\t%.17g\n", mesh[3 * (size_t)i], mesh[3 * (size_t)i + 1],
                     mesh[3 * (size_t)i + 2], sx[i], sy[i], sz[i]);
    if (std::ferror(f.get())) IO_FAIL(LLG_E_INVALID, "write error on %s", path);
    return 0;
}

int llg_write_vtu(const char* path, uint32_t n, const double* pos, const double* sx, const double* sy,
                  const double* sz) {
    if (!path || !pos || !sx || !sy || !sz) IO_FAIL(LLG_E_INVALID, "null argument");
    std::unique_ptr<FILE, FileCloser> f(std::fopen(path, "w"));
    if (!f) IO_FAIL(LLG_E_INVALID, "could not open %s for writing", path);
    FILE* o = f.get();
    std::fprintf(o, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\">\n<UnstructuredGrid>\n");
    std::fprintf(o, "<Piece NumberOfPoints=\"%u\"  NumberOfCells=\"1\">\n<PointData Scalar=\"Spin\">\n", n);
    std::fprintf(o, "<DataArray type=\"Float32\" Name=\"Spin\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    for (uint32_t i = 0; i < n; ++i) std::fprintf(o, "%.9g\t%.9g\t%.9g\n", sx[i], sy[i], sz[i]);
    std::fprintf(o, "</DataArray>\n</PointData>\n<CellData>\n</CellData>\n<Points>\n");
    std::fprintf(o, "<DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    for (uint32_t i = 0; i < n; ++i)
        std::fprintf(o, "%.9g\t%.9g\t%.9g\n", pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]);
    std::fprintf(o, "</DataArray>\n</Points>\n<Cells>\n");
    std::fprintf(o, "<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n1\n</DataArray>\n");
    std::fprintf(o, "<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n1\n</DataArray>\n");
    std::fprintf(o, "<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n1\n</DataArray>\n");
    std::fprintf(o, "</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
    if (std::ferror(o)) IO_FAIL(LLG_E_INVALID, "write error on %s", path);
    return 0;
}

}  // extern "C"
