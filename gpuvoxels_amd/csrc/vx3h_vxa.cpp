// vx3h_vxa.cpp -- VXA/VXD loader and report writer behind the C ABI.
//
// Input side of the drop-in boundary: a base VXA text, optionally patched by a VXD
// whose children carry replace="VXA.a.b" (Voxelyze3/src/Utils/ctool.h:49-57), is
// read the way VX3_SimulationManager::readVXD reads it
// (Voxelyze3/src/VX3/VX3_SimulationManager.cu:277-380):
//   * CVX_Sim::ReadVXA + Import (Voxelyze3/src/VXA/VX_Sim.cpp:59-294), walking the
//     document with the cursor rules of CXML_Rip (Voxelyze3/src/old/XML_Rip.cpp:165-236):
//     FindElement looks among the children of the current element, or -- when the
//     current element has the searched name -- among its following siblings;
//     FindLoadElement reads the element's first text child with whitespace condensed;
//   * the VX3-only simulator keys through property-tree paths ("VXA.Simulator...."),
//     whose values are the concatenated raw text of the element;
//   * MathTree programs as the reversed breadth-first order of the element subtree
//     (VX3_SimulationManager.cu:157-275).
// The result is a vx3h_sim_desc: palette values before Voxelyze's setters, voxels in
// structure-index order, phase offsets, boundary conditions resolved to per-voxel
// externals (Import's loop, VX_Sim.cpp:108-141), environment and simulator switches.
//
// Output side: the report XML of vx3_node_worker
// (Voxelyze3/src/Executables/vx3_node_worker.cu:98-141), entries sorted by fitness
// descending with NaN last (Voxelyze3/src/VX3/VX3_SimulationResult.h:26-33).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/vx3h.h"
#include "vx3h_internal.h"

using vx3::g_err;

namespace {

// ------------------------------------------------------------------ XML document
struct XNode {
    std::string name;                                        // element name ("<xmlcomment>" for a comment)
    std::vector<std::pair<std::string, std::string>> attrs;
    std::string data;                                        // concatenated raw text / CDATA of direct children
    std::vector<XNode> kids;                                 // element and comment children in document order
};

struct XParser {
    const char* p;
    const char* end;
    std::string err;

    bool fail(const std::string& m) { if (err.empty()) err = m; return false; }
    void skipWs() { while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) p++; }
    bool starts(const char* s) const { size_t n = strlen(s); return (size_t)(end - p) >= n && memcmp(p, s, n) == 0; }

    static void appendUtf8(std::string& o, unsigned long c) {
        if (c < 0x80) o += (char)c;
        else if (c < 0x800) { o += (char)(0xC0 | (c >> 6)); o += (char)(0x80 | (c & 0x3F)); }
        else if (c < 0x10000) { o += (char)(0xE0 | (c >> 12)); o += (char)(0x80 | ((c >> 6) & 0x3F)); o += (char)(0x80 | (c & 0x3F)); }
        else { o += (char)(0xF0 | (c >> 18)); o += (char)(0x80 | ((c >> 12) & 0x3F)); o += (char)(0x80 | ((c >> 6) & 0x3F)); o += (char)(0x80 | (c & 0x3F)); }
    }
    std::string decode(const char* a, const char* b) {
        std::string o;
        while (a < b) {
            if (*a != '&') { o += *a++; continue; }
            const char* semi = (const char*)memchr(a, ';', b - a);
            if (!semi) { o += *a++; continue; }
            std::string ent(a + 1, semi);
            if (ent == "lt") o += '<';
            else if (ent == "gt") o += '>';
            else if (ent == "amp") o += '&';
            else if (ent == "quot") o += '"';
            else if (ent == "apos") o += '\'';
            else if (!ent.empty() && ent[0] == '#') {
                unsigned long c = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X') ? strtoul(ent.c_str() + 2, nullptr, 16)
                                                                                      : strtoul(ent.c_str() + 1, nullptr, 10);
                appendUtf8(o, c);
            } else { o.append(a, semi + 1); }
            a = semi + 1;
        }
        return o;
    }
    bool name(std::string& out) {
        const char* s = p;
        while (p < end && !(*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n' || *p == '>' || *p == '/' || *p == '=' || *p == '<')) p++;
        if (p == s) return fail("expected a name");
        out.assign(s, p);
        return true;
    }
    bool skipMisc() {   // declarations, comments, doctype, whitespace between top-level nodes
        for (;;) {
            skipWs();
            if (starts("<?")) { const char* q = strstr(p, "?>"); if (!q || q >= end) return fail("unterminated <?"); p = q + 2; }
            else if (starts("<!--")) { const char* q = strstr(p + 4, "-->"); if (!q || q >= end) return fail("unterminated comment"); p = q + 3; }
            else if (starts("<!DOCTYPE")) { const char* q = (const char*)memchr(p, '>', end - p); if (!q) return fail("bad doctype"); p = q + 1; }
            else return true;
        }
    }
    bool element(XNode& n, int depth) {
        if (depth > 256) return fail("nesting too deep");
        if (p >= end || *p != '<') return fail("expected '<'");
        p++;
        if (!name(n.name)) return false;
        for (;;) {
            skipWs();
            if (p >= end) return fail("unterminated start tag <" + n.name + ">");
            if (*p == '/') {
                if (p + 1 < end && p[1] == '>') { p += 2; return true; }
                return fail("bad empty-element tag");
            }
            if (*p == '>') { p++; break; }
            std::string an;
            if (!name(an)) return false;
            skipWs();
            if (p >= end || *p != '=') return fail("attribute without value in <" + n.name + ">");
            p++;
            skipWs();
            if (p >= end || (*p != '"' && *p != '\'')) return fail("unquoted attribute in <" + n.name + ">");
            const char q = *p++;
            const char* s = p;
            while (p < end && *p != q) p++;
            if (p >= end) return fail("unterminated attribute");
            n.attrs.push_back({an, decode(s, p)});
            p++;
        }
        for (;;) {
            if (p >= end) return fail("unterminated element <" + n.name + ">");
            if (starts("</")) {
                p += 2;
                std::string en;
                if (!name(en)) return false;
                if (en != n.name) return fail("mismatched </" + en + "> for <" + n.name + ">");
                skipWs();
                if (p >= end || *p != '>') return fail("bad end tag");
                p++;
                return true;
            }
            if (starts("<!--")) {
                const char* q = strstr(p + 4, "-->");
                if (!q || q >= end) return fail("unterminated comment");
                n.kids.emplace_back();
                n.kids.back().name = "<xmlcomment>";
                n.kids.back().data.assign(p + 4, q);
                p = q + 3;
                continue;
            }
            if (starts("<![CDATA[")) {
                const char* s = p + 9;
                const char* q = strstr(s, "]]>");
                if (!q || q >= end) return fail("unterminated CDATA");
                n.data.append(s, q);
                p = q + 3;
                continue;
            }
            if (starts("<?")) { const char* q = strstr(p, "?>"); if (!q || q >= end) return fail("unterminated <?"); p = q + 2; continue; }
            if (*p == '<') {
                n.kids.emplace_back();
                if (!element(n.kids.back(), depth + 1)) return false;
                continue;
            }
            const char* s = p;
            while (p < end && *p != '<') p++;
            n.data += decode(s, p);
        }
    }
    bool document(XNode& root) {
        if (!skipMisc()) return false;
        if (!element(root, 0)) return false;
        return skipMisc();
    }
};

bool parseXml(const char* text, XNode& root, std::string& err) {
    XParser x{text, text + strlen(text), ""};
    if (!x.document(root)) { err = "XML: " + x.err; return false; }
    return true;
}

// Whitespace-condensed text: leading and trailing whitespace dropped, inner runs -> one space.
std::string condense(const std::string& s) {
    std::string o;
    bool ws = false;
    for (char c : s) {
        if (c == ' ' || c == '\t' || c == '\r' || c == '\n') { ws = !o.empty(); continue; }
        if (ws) { o += ' '; ws = false; }
        o += c;
    }
    return o;
}

std::vector<std::string> splitPath(const std::string& path) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : path) { if (c == '.') { out.push_back(cur); cur.clear(); } else cur += c; }
    out.push_back(cur);
    return out;
}

// ptree_merge (ctool.h:49-57): every child of <VXD> (or of the document, when its root is
// not VXD) with a replace attribute replaces the first node at that path, creating the
// path when missing; the replaced node keeps its name and takes the child's content.
void mergeVxd(XNode& vxa, const XNode& vxdRoot) {
    std::vector<const XNode*> items;
    if (vxdRoot.name == "VXD") for (const XNode& k : vxdRoot.kids) items.push_back(&k);
    else items.push_back(&vxdRoot);
    for (const XNode* k : items) {
        std::string rep;
        for (auto& a : k->attrs) if (a.first == "replace") rep = a.second;
        if (rep.empty()) continue;
        std::vector<std::string> comps = splitPath(rep);
        if (comps.empty() || comps[0] != vxa.name) continue;   // would land outside the document root
        XNode* cur = &vxa;
        for (size_t i = 1; i < comps.size(); i++) {
            XNode* nx = nullptr;
            for (XNode& c : cur->kids) if (c.name == comps[i]) { nx = &c; break; }
            if (!nx) { cur->kids.emplace_back(); cur->kids.back().name = comps[i]; nx = &cur->kids.back(); }
            cur = nx;
        }
        if (cur == &vxa) continue;
        const std::string nm = cur->name;
        *cur = *k;
        cur->name = nm;
    }
}

// property-tree style lookup: first child matching each component; value = raw data
const XNode* ptFind(const XNode& root, const std::string& path) {
    std::vector<std::string> comps = splitPath(path);
    if (comps.empty() || comps[0] != root.name) return nullptr;
    const XNode* cur = &root;
    for (size_t i = 1; i < comps.size(); i++) {
        const XNode* nx = nullptr;
        for (const XNode& c : cur->kids) if (c.name == comps[i]) { nx = &c; break; }
        if (!nx) return nullptr;
        cur = nx;
    }
    return cur;
}
template <typename T> T ptGet(const XNode& root, const std::string& path, T def) {
    const XNode* n = ptFind(root, path);
    if (!n) return def;
    std::istringstream is(n->data);
    T v;
    is >> v;
    if (is.fail()) return def;
    is >> std::ws;
    return is.eof() ? v : def;
}
template <> bool ptGet<bool>(const XNode& root, const std::string& path, bool def) {
    const XNode* n = ptFind(root, path);
    if (!n) return def;
    std::istringstream is(n->data);
    bool v;
    is >> v;
    if (is.fail()) { is.clear(); is.str(n->data); is >> std::boolalpha >> v; }
    if (is.fail()) return def;
    is >> std::ws;
    return is.eof() ? v : def;
}

// ------------------------------------------------------------------ CXML_Rip cursor
struct Rip {
    struct Frame { const XNode* parent; int idx; const XNode* el; std::string tag; };
    std::vector<Frame> st;
    explicit Rip(const XNode& root) { st.push_back({nullptr, 0, &root, "Root"}); }
    const XNode* cur() const { return st.back().el; }
    void up() { if (st.size() > 1) st.pop_back(); }
    bool find(const std::string& tag) {
        Frame& f = st.back();
        if (tag == f.tag && f.parent) {
            for (int i = f.idx + 1; i < (int)f.parent->kids.size(); i++)
                if (f.parent->kids[i].name == tag) { f.idx = i; f.el = &f.parent->kids[i]; return true; }
            up();
            return false;
        }
        const XNode* c = f.el;
        for (int i = 0; i < (int)c->kids.size(); i++)
            if (c->kids[i].name == tag) { st.push_back({c, i, &c->kids[i], tag}); return true; }
        return false;
    }
    // FindLoadElement: the element stays entered when it has no text (as the reference's does)
    bool load(const std::string& tag, std::string& out, bool stayDown = false) {
        if (!find(tag)) return false;
        std::string t = condense(cur()->data);
        if (t.empty()) return false;
        out = t;
        if (!stayDown) up();
        return true;
    }
    bool load(const std::string& tag, double& v, bool stayDown = false) {
        std::string s;
        if (!load(tag, s, stayDown)) return false;
        v = atof(s.c_str());
        return true;
    }
    bool load(const std::string& tag, float& v, bool stayDown = false) {
        double d;
        if (!load(tag, d, stayDown)) return false;
        v = (float)d;
        return true;
    }
    bool load(const std::string& tag, int& v, bool stayDown = false) {
        std::string s;
        if (!load(tag, s, stayDown)) return false;
        v = atoi(s.c_str());
        return true;
    }
    bool load(const std::string& tag, bool& v, bool stayDown = false) {
        int i;
        if (!load(tag, i, stayDown)) return false;
        v = i != 0;
        return true;
    }
    std::string attr(const std::string& a) const {
        for (auto& kv : cur()->attrs) if (kv.first == a) return kv.second;
        return "";
    }
};

// ------------------------------------------------------------------ VXA content
struct Region {                 // CVX_FRegion (Voxelyze3/src/old/VX_FRegion.cpp:165-215)
    int prim = -1;              // 0 box, 1 cylinder, 2 sphere, 3 mesh
    double X = 0, Y = 0, Z = 0, dX = 0, dY = 0, dZ = 0, radius = 0;
    int dof = 0;
    double force[3] = {0, 0, 0}, torque[3] = {0, 0, 0}, disp[3] = {0, 0, 0}, angDisp[3] = {0, 0, 0};
};

struct PalMat {                 // CVXC_Material (VX_Object.cpp:1373-1460), SINGLE materials
    int type = 0;               // MatMode: 0 SINGLE, 1 INTERNAL, 2 EXTERNAL, 3 DITHER
    int model = 0;
    std::vector<double> strain, stress;
    int isMeasured = 1;
    double thermalOnAfter = 0;
    int fixed = 0;
    double E = 0, plasticMod = 0, yieldStress = 0, failStress = 0, density = 0, nu = 0, cte = 0, uS = 0, uD = 0;
    float rgba[4] = {0.5f, 0.5f, 0.5f, 1.0f};   // Display colour (CVXC_Material::ClearMaterial defaults)
    double cilia = 0, ciliaOn = 0;              // Cilia, TurnOnCiliaAfterThisManySeconds
    double removeAfter = 0;                     // RemoveFromSimulationAfterThisManySeconds
    int isPaceMaker = 0;                        // signals (VX_Object.cpp:1420-1425)
    double paceMakerPeriod = 0, signalValueDecay = 0.9, signalTimeDelay = 0.03, inactivePeriod = 0.03;
    int isTarget = 0;
    int sticky = 0;
};

bool readRegion(Rip& r, Region& g) {
    if (!r.load("PrimType", g.prim)) return false;
    r.load("X", g.X); r.load("Y", g.Y); r.load("Z", g.Z);
    r.load("dX", g.dX); r.load("dY", g.dY); r.load("dZ", g.dZ);
    r.load("Radius", g.radius);
    int dof;
    if (r.load("DofFixed", dof)) g.dof = (int)(char)dof & 0x3F;
    else { bool oldFixed; g.dof = r.load("Fixed", oldFixed) ? (oldFixed ? 0x3F : 0) : 0; }
    const char* fx[3] = {"ForceX", "ForceY", "ForceZ"};
    const char* tq[3] = {"TorqueX", "TorqueY", "TorqueZ"};
    const char* ds[3] = {"DisplaceX", "DisplaceY", "DisplaceZ"};
    const char* ad[3] = {"AngDisplaceX", "AngDisplaceY", "AngDisplaceZ"};
    for (int k = 0; k < 3; k++) {
        r.load(fx[k], g.force[k]); r.load(tq[k], g.torque[k]);
        r.load(ds[k], g.disp[k]); r.load(ad[k], g.angDisp[k]);
    }
    return true;
}

// IsTouching(P, Dist vector, Envelope) of the region primitives (VX_FRegion.cpp:420-430,526-604)
bool touching(const Region& g, const double P[3], const double D[3], const double W[3]) {
    const double wmax = std::max(W[0], std::max(W[1], W[2]));
    if (g.prim == 0) {
        double Ps[3], Ds[3];
        for (int k = 0; k < 3; k++) { Ps[k] = P[k] / W[k]; Ds[k] = D[k] / W[k]; }
        return Ps[0] + Ds[0] > g.X && Ps[0] - Ds[0] < g.X + g.dX && Ps[1] + Ds[1] > g.Y && Ps[1] - Ds[1] < g.Y + g.dY &&
               Ps[2] + Ds[2] > g.Z && Ps[2] - Ds[2] < g.Z + g.dZ;
    }
    const double pos[3] = {g.X * W[0], g.Y * W[1], g.Z * W[2]};
    const double size[3] = {g.dX * W[0], g.dY * W[1], g.dZ * W[2]};
    const double rad = g.radius * wmax;
    if (g.prim == 2) {
        const double dist = sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
        return sqrt((P[0] - pos[0]) * (P[0] - pos[0]) + (P[1] - pos[1]) * (P[1] - pos[1]) + (P[2] - pos[2]) * (P[2] - pos[2])) < rad + dist;
    }
    if (g.prim == 1) {   // CP_Cylinder::IsIn
        if (g.dX != 0) return P[0] > pos[0] && P[0] < pos[0] + size[0] && sqrt((P[1] - pos[1]) * (P[1] - pos[1]) + (P[2] - pos[2]) * (P[2] - pos[2])) < rad;
        if (g.dY != 0) return P[1] > pos[1] && P[1] < pos[1] + size[1] && sqrt((P[0] - pos[0]) * (P[0] - pos[0]) + (P[2] - pos[2]) * (P[2] - pos[2])) < rad;
        if (g.dZ != 0) return P[2] > pos[2] && P[2] < pos[2] + size[2] && sqrt((P[0] - pos[0]) * (P[0] - pos[0]) + (P[1] - pos[1]) * (P[1] - pos[1])) < rad;
        return false;
    }
    return false;
}

std::string fromBase64(const std::string& s) {   // standard alphabet, stops at '=' or a foreign character
    static const std::string A = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string out;
    unsigned v = 0;
    int bits = 0;
    for (char c : s) {
        size_t k = A.find(c);
        if (c == '=' || k == std::string::npos) break;
        v = (v << 6) | (unsigned)k;
        bits += 6;
        if (bits >= 8) { bits -= 8; out += (char)((v >> bits) & 0xFF); }
    }
    return out;
}

std::vector<std::string> splitComma(const std::string& s) {
    std::vector<std::string> out;
    std::istringstream is(s);
    std::string item;
    while (std::getline(is, item, ',')) out.push_back(item);
    return out;
}

// MathTree: reversed breadth-first order of the subtree, then mtEND
// (VX3_SimulationManager.cu:157-275).  The subtree is walked as a property tree:
// an element's attributes form a first child "<xmlattr>" (one child per attribute)
// and comments are "<xmlcomment>" children.  Tokens are taken from the end of the
// breadth-first order; at the first token the reference cannot convert it stops,
// leaving an mtVAR with value 0 for an unknown variable name, an mtCONST 0 for a
// constant without a number, and nothing for an unknown operation, and the
// program ends there (the token array's default is mtEND).
struct MtNode { std::string op, value; std::vector<MtNode> kids; };

MtNode mtTree(const XNode& x) {
    MtNode m;
    m.op = x.name;
    m.value = x.data;
    if (!x.attrs.empty()) {
        MtNode a;
        a.op = "<xmlattr>";
        for (auto& kv : x.attrs) { MtNode c; c.op = kv.first; c.value = kv.second; a.kids.push_back(c); }
        m.kids.push_back(a);
    }
    for (const XNode& k : x.kids) m.kids.push_back(mtTree(k));
    return m;
}

bool parseMathTree(const XNode& root, const std::string& path, std::vector<vx3h_token>& out, std::string& err) {
    out.clear();
    const XNode* n = ptFind(root, path);
    if (!n) return true;
    static const char* ops[] = {"mtEND", "mtCONST", "mtE", "mtPI", "mtVAR", "mtADD", "mtSUB", "mtMUL", "mtDIV", "mtPOW",
                                "mtSQRT", "mtSIN", "mtCOS", "mtTAN", "mtATAN", "mtLOG", "mtINT", "mtABS", "mtNOT",
                                "mtGREATERTHAN", "mtLESSTHAN", "mtAND", "mtOR", "mtNORMALCDF"};
    static const char* vars[] = {"x", "y", "z", "hit", "t", "angle", "targetCloseness", "numClosePairs", "num_voxel"};
    const MtNode tree = mtTree(*n);
    std::vector<const MtNode*> bfs;
    std::vector<const MtNode*> frontier{&tree};
    for (size_t h = 0; h < frontier.size(); h++)
        for (const MtNode& c : frontier[h]->kids) { bfs.push_back(&c); frontier.push_back(&c); }
    auto rtrim = [](std::string v) {
        while (!v.empty() && (v.back() == ' ' || v.back() == '\t' || v.back() == '\n' || v.back() == '\r')) v.pop_back();
        return v;
    };
    for (auto it = bfs.rbegin(); it != bfs.rend(); ++it) {
        const std::string op = rtrim((*it)->op), val = rtrim((*it)->value);
        int code = -1;
        for (int i = 0; i < (int)(sizeof(ops) / sizeof(ops[0])); i++) if (op == ops[i]) code = i;
        if (code < 0) break;                        // "Token Operation not implemented"
        vx3h_token tk{code, 0.0};
        if (code == 4) {
            int k = -1;
            for (int i = 0; i < 9; i++) if (val == vars[i]) k = i;
            if (k < 0) { out.push_back(tk); break; }   // "No such variable": the slot keeps mtVAR 0
            tk.value = k;
        } else if (code == 1) {
            char* e = nullptr;
            const char* s = val.c_str();
            const double v = strtod(s, &e);
            if (e == s) { out.push_back(tk); break; }  // "mtCONST with no number": mtCONST 0
            tk.value = v;
        }
        out.push_back(tk);
        if (code == 0) break;
    }
    if (out.empty() || out.back().op != 0) out.push_back({0, 0.0});
    if (out.size() > VX3H_MAX_TOKENS) { err = path + ": too many MathTree tokens"; return false; }
    return true;
}

}  // namespace

struct vx3h_vxa {
    vx3h_sim_desc desc;
    std::vector<vx3h_material_desc> mats;
    std::vector<int> ijk, mat;
    std::vector<double> phase;
    std::vector<vx3h_external> exts;
    std::vector<vx3h_token> stop, fit, ff[3], att[5];
    std::vector<double> baseCilia, shiftCilia;
    std::string rawPrint;
    vx3h_vxa_info info;
};

namespace {

vx3h_status readVxa(const XNode& doc, vx3h_vxa& o, std::string& err) {
    Rip r(doc);
    // ---- Simulator (CVX_Sim::ReadXML, VX_Sim.cpp:185-294)
    double dtFrac = 0.9;
    float bondZ = 0.1f, colZ = 1.0f, slowZ = 1.0f;   // values the reader applies for absent <Damping> tags
    bool volumeEffects = false;
    if (r.find("Simulator")) {
        if (r.find("Integration")) {
            if (!r.load("DtFrac", dtFrac)) dtFrac = 0.9;
            r.up();
        }
        if (r.find("Damping")) {
            float t;
            bondZ = r.load("BondDampingZ", t) ? t : 0.1f;
            colZ = r.load("ColDampingZ", t) ? t : 1.0f;
            slowZ = r.load("SlowDampingZ", t) ? t : 1.0f;
            r.up();
        }
        if (r.find("Features")) {
            if (!r.load("VolumeEffectsEnabled", volumeEffects)) volumeEffects = false;
            r.up();
        }
        r.up();
    }
    // ---- Environment (CVX_Environment::ReadXML, VX_Environment.cpp:101-180)
    std::vector<Region> bcs;
    bool gravEnabled = true, floorEnabled = true, tempEnabled = true, varyTemp = false;
    double gravAcc = -9.81, tempBase = 25, tempAmp = 0, tempPeriod = 0.1;
    if (r.find("Environment")) {
        auto readRegions = [&](const char* group, const char* countTag) -> bool {
            if (!r.find(group)) return false;
            int n = 0;
            if (!r.load(countTag, n)) n = 0;
            for (int i = 0; i < n; i++) {
                if (!r.find("FRegion")) { err = std::string(group) + ": fewer FRegion elements than declared"; return true; }
                Region g;
                if (!readRegion(r, g)) g.prim = -1;
                bcs.push_back(g);
            }
            if (n != 0) r.up();
            r.up();
            return true;
        };
        if (!readRegions("Boundary_Conditions", "NumBCs")) {
            readRegions("Fixed_Regions", "NumFixed");
            readRegions("Forced_Regions", "NumForced");
        }
        if (!err.empty()) return VX3H_EPARSE;
        if (r.find("Gravity")) {
            if (!r.load("GravEnabled", gravEnabled)) gravEnabled = false;
            if (!r.load("GravAcc", gravAcc)) gravAcc = -9.81;
            if (!r.load("FloorEnabled", floorEnabled)) floorEnabled = false;
            r.up();
        }
        if (r.find("Thermal")) {
            if (!r.load("TempEnabled", tempEnabled)) tempEnabled = false;
            if (!r.load("TempBase", tempBase)) tempBase = 25;
            if (!r.load("TempAmplitude", tempAmp)) {
                double legacy;
                tempAmp = r.load("TempAmp", legacy) ? legacy - tempBase : 0;
            }
            if (!r.load("VaryTempEnabled", varyTemp)) varyTemp = false;
            if (!r.load("TempPeriod", tempPeriod)) tempPeriod = 0.1;
            r.up();
        }
        r.up();
    }
    (void)gravEnabled;   // Import overrides the gravity switch with -GravAcc/9.80665 (VX_Sim.cpp:65-66,485)
    // ---- VXC (CVX_Object::ReadXML, VX_Object.cpp:446-471)
    double latDim = 0.001, dimAdj[3] = {1, 1, 1}, lineOff[2] = {0, 0}, layerOff[2] = {0, 0};
    std::vector<PalMat> pal;
    int XV = 1, YV = 1, ZV = 1;
    std::vector<unsigned char> cells;
    std::vector<double> cellPhase;
    std::vector<double> cellCilia[2];   // BaseCiliaForce, ShiftCiliaForce: 3 per occupied cell
    bool havePhase = false;
    if (r.find("VXC") || r.find("DMF")) {
        const std::string version = r.attr("Version");
        if (r.find("Lattice")) {   // CVXC_Lattice::ReadXML (VX_Object.cpp:1076-1086)
            if (!r.load("Lattice_Dim", latDim)) latDim = 0.001;
            if (!r.load("X_Dim_Adj", dimAdj[0])) dimAdj[0] = 1.0;
            if (!r.load("Y_Dim_Adj", dimAdj[1])) dimAdj[1] = 1.0;
            if (!r.load("Z_Dim_Adj", dimAdj[2])) dimAdj[2] = 1.0;
            if (!r.load("X_Line_Offset", lineOff[0])) lineOff[0] = 0.0;
            if (!r.load("Y_Line_Offset", lineOff[1])) lineOff[1] = 0.0;
            if (!r.load("X_Layer_Offset", layerOff[0])) layerOff[0] = 0.0;
            if (!r.load("Y_Layer_Offset", layerOff[1])) layerOff[1] = 0.0;
            r.up();
        }
        if (r.find("Palette")) {
            // one material object is read over and over (ReadPaletteXML, VX_Object.cpp:477-488):
            // a Display or Mechanical block that a material lacks keeps the values of the
            // material before it
            PalMat m;
            while (r.find("Material")) {
                m.strain.clear();
                m.stress.clear();
                int mt;
                if (r.load("MatType", mt)) m.type = mt;
                else {   // the reference's detective work, including its fall-through to SINGLE
                    if (r.find("Structure")) { r.up(); m.type = 1; }
                    if (r.find("RandIndex1")) { r.up(); m.type = 3; }
                    else if (r.find("VXC") || r.find("DMF")) { r.up(); m.type = 2; }
                    else m.type = 0;
                }
                std::string name;
                r.load("Name", name);   // moves the cursor as the reference's read does
                if (r.find("Display")) {
                    if (!r.load("Red", m.rgba[0])) m.rgba[0] = 0.5f;
                    if (!r.load("Green", m.rgba[1])) m.rgba[1] = 0.5f;
                    if (!r.load("Blue", m.rgba[2])) m.rgba[2] = 0.5f;
                    if (!r.load("Alpha", m.rgba[3])) m.rgba[3] = 1.0f;
                    r.up();
                }
                if (m.type == 0 && r.find("Mechanical")) {
                    if (!r.load("MatModel", m.model)) m.model = 0;
                    if (r.find("SSData")) {
                        int np = 0;
                        if (!r.load("NumDataPts", np)) np = 0;
                        auto series = [&](const char* group, const char* item, std::vector<double>& v) {
                            if (!r.find(group)) return;
                            double last = 0;
                            for (int i = 0; i < np; i++) {
                                if (r.load(item, last, true)) v.push_back(last);
                                else { err = std::string(group) + ": fewer points than NumDataPts"; return; }
                            }
                            r.up();
                            r.up();
                        };
                        series("StrainData", "Strain", m.strain);
                        series("StressData", "Stress", m.stress);
                        r.up();
                        if (!err.empty()) return VX3H_EPARSE;
                    }
                    if (!r.load("isMeasured", m.isMeasured)) m.isMeasured = 1;
                    if (!r.load("TurnOnThermalExpansionAfterThisManySeconds", m.thermalOnAfter)) m.thermalOnAfter = 0;
                    if (!r.load("Fixed", m.fixed)) m.fixed = 0;
                    if (!r.load("TurnOnCiliaAfterThisManySeconds", m.ciliaOn)) m.ciliaOn = 0.0;
                    if (!r.load("RemoveFromSimulationAfterThisManySeconds", m.removeAfter)) m.removeAfter = 0.0;
                    if (!r.load("isPaceMaker", m.isPaceMaker)) m.isPaceMaker = 0;
                    if (!r.load("PaceMakerPeriod", m.paceMakerPeriod)) m.paceMakerPeriod = 0;
                    if (!r.load("signalValueDecay", m.signalValueDecay)) m.signalValueDecay = 0.9;
                    if (!r.load("signalTimeDelay", m.signalTimeDelay)) m.signalTimeDelay = 0.03;
                    if (!r.load("inactivePeriod", m.inactivePeriod)) m.inactivePeriod = 0.03;
                    if (!r.load("isTarget", m.isTarget)) m.isTarget = 0;
                    if (!r.load("Sticky", m.sticky)) m.sticky = 0;   // VXA/VX_Object.cpp:1436
                    if (!r.load("Cilia", m.cilia)) m.cilia = 0;
                    if (!r.load("Elastic_Mod", m.E)) m.E = 0;
                    if (!r.load("Plastic_Mod", m.plasticMod)) m.plasticMod = 0;
                    if (!r.load("Yield_Stress", m.yieldStress)) m.yieldStress = 0;
                    if (!r.load("Fail_Stress", m.failStress)) m.failStress = 0;
                    if (!r.load("Density", m.density)) m.density = 0;
                    if (!r.load("Poissons_Ratio", m.nu)) m.nu = 0;
                    if (!r.load("CTE", m.cte)) m.cte = 0;
                    if (!r.load("uStatic", m.uS)) m.uS = 0;
                    if (!r.load("uDynamic", m.uD)) m.uD = 0;
                    r.up();
                }
                pal.push_back(m);
            }
            r.up();
        }
        if (r.find("Structure")) {   // CVXC_Structure::ReadXML (VX_Object.cpp:1760-1960)
            const std::string comp = r.attr("Compression");
            if (!r.load("X_Voxels", XV)) XV = 1;
            if (!r.load("Y_Voxels", YV)) YV = 1;
            if (!r.load("Z_Voxels", ZV)) ZV = 1;
            if (XV <= 0 || YV <= 0 || ZV <= 0 || (long)XV * YV * ZV > (1L << 28)) { err = "bad structure dimensions"; return VX3H_EPARSE; }
            cells.assign((size_t)XV * YV * ZV, 0);
            const size_t layer = (size_t)XV * YV;
            if (version == "" || version == "0.93" || version == "0.94") {
                if (!r.find("Data")) { err = "Structure without Data"; return VX3H_EPARSE; }
                for (int z = 0; z < ZV; z++) {
                    std::string raw, in;
                    if (!r.load("Layer", raw, true)) { err = "missing structure layer"; return VX3H_EPARSE; }
                    if (comp == "ASCII_READABLE") { in = raw; for (char& c : in) c = (char)(c - 48); }
                    else if (comp == "RAW_DATA") in = raw;
                    else if (comp == "ZLIB" || comp == "QT_ZLIB") { err = comp + " compressed VXC is not supported"; return VX3H_EUNSUPPORTED; }
                    else in = fromBase64(raw);
                    if (in.size() != layer) { err = "Voxel layer data not present or does not match expected size."; return VX3H_EPARSE; }
                    memcpy(&cells[z * layer], in.data(), layer);
                }
                r.up();
                r.up();
            } else if (version == "0.92") {
                std::string raw;
                r.load("Data", raw);
                size_t it = 0;
                for (size_t k = 0; k < cells.size() && it < raw.size(); it++)
                    if (raw[it] != '\n' && raw[it] != '\r') cells[k++] = (unsigned char)raw[it];
            } else { err = "unknown VXC version " + version; return VX3H_EUNSUPPORTED; }
            if (r.find("PhaseOffset")) {
                havePhase = true;
                for (int z = 0; z < ZV; z++) {
                    std::string raw;
                    if (!r.load("Layer", raw, true)) { err = "missing PhaseOffset layer"; return VX3H_EPARSE; }
                    std::vector<std::string> vals = splitComma(raw);
                    for (size_t k = 0; k < layer; k++) {
                        if ((signed char)cells[z * layer + k] > 0) {
                            if (k >= vals.size()) { err = "Data in PhaseOffset.Layers is too short"; return VX3H_EPARSE; }
                            cellPhase.push_back(atof(vals[k].c_str()));
                        }
                    }
                }
                r.up();
                r.up();
            }
            // CVXC_Structure::ReadXML (VX_Object.cpp:1886-1958): three values per cell k of a
            // layer at 3k..3k+2, for the occupied cells; a short layer pads with zeros
            const char* ciliaTags[2] = {"BaseCiliaForce", "ShiftCiliaForce"};
            for (int c = 0; c < 2; c++) {
                if (!r.find(ciliaTags[c])) continue;
                for (int z = 0; z < ZV; z++) {
                    std::string raw;
                    r.load("Layer", raw, true);
                    std::vector<std::string> vals = splitComma(raw);
                    for (size_t k = 0; k < layer; k++) {
                        if ((signed char)cells[z * layer + k] > 0)
                            for (int j = 0; j < 3; j++)
                                cellCilia[c].push_back(3 * k + j < vals.size() ? atof(vals[3 * k + j].c_str()) : 0.0);
                    }
                }
                r.up();
                r.up();
            }
            r.up();
        }
        r.up();
    }

    // ---- Import (VX_Sim.cpp:59-153)
    std::vector<int> palToMat(pal.size() + 1, -1);
    for (size_t i = 0; i < pal.size(); i++) {
        const PalMat& p = pal[i];
        if (p.type != 0) { err = "only SINGLE palette materials are supported"; return VX3H_EUNSUPPORTED; }
        if (p.model < 0 || p.model > 3) { err = "unknown MatModel"; return VX3H_EPARSE; }
        vx3h_material_desc d;
        memset(&d, 0, sizeof(d));
        d.mat_model = p.model;
        d.elastic_mod = p.E;
        d.plastic_mod = p.plasticMod;
        d.yield_stress = p.yieldStress;
        d.fail_stress = p.failStress;
        d.density = p.density;
        d.poissons_ratio = p.nu;
        d.cte = p.cte;
        d.u_static = p.uS;
        d.u_dynamic = p.uD;
        d.zeta_internal = bondZ;
        d.zeta_global = slowZ;
        d.zeta_collision = colZ;
        if (p.model == 3) {
            if (p.strain.size() != p.stress.size() || p.strain.size() < 2 || p.strain.size() > VX3H_MAX_CURVE) {
                err = "stress/strain data must have 2.." + std::to_string(VX3H_MAX_CURVE) + " points of equal count";
                return VX3H_EUNSUPPORTED;
            }
            d.n_data = (int)p.strain.size();
            for (int k = 0; k < d.n_data; k++) { d.strain_data[k] = p.strain[k]; d.stress_data[k] = p.stress[k]; }
        }
        d.is_measured = p.isMeasured;
        d.fixed = p.fixed != 0;
        d.thermal_on_after = p.thermalOnAfter;
        d.cilia = p.cilia;
        d.cilia_on_after = p.ciliaOn;
        d.remove_after = p.removeAfter;
        d.is_pacemaker = p.isPaceMaker != 0;
        d.pacemaker_period = p.paceMakerPeriod;
        d.signal_value_decay = p.signalValueDecay;
        d.signal_time_delay = p.signalTimeDelay;
        d.inactive_period = p.inactivePeriod;
        d.is_target = p.isTarget != 0;
        d.sticky = p.sticky != 0;
        d.matid = (int)i + 1;   // palette index; 0 is the erase material
        for (int k = 0; k < 4; k++) {   // GetRedi() etc. (VX_Object.h:419-425), then setRed's clamp
            const int c = (int)(p.rgba[k] * 255);
            d.color[k] = c > 255 ? 255 : (c < 0 ? 0 : c);
        }
        palToMat[i + 1] = (int)o.mats.size();
        o.mats.push_back(d);
    }
    const double vs = latDim * dimAdj[0];     // CVoxelyze voxel size: GetLatDimEnv().x
    const double latEnv[3] = {latDim * dimAdj[0], latDim * dimAdj[1], latDim * dimAdj[2]};
    const size_t layer = (size_t)XV * YV;
    std::vector<size_t> voxCell;
    for (size_t i = 0; i < cells.size(); i++) {
        const int m = (int)(signed char)cells[i];
        if (m - 1 < 0) continue;
        if (m >= (int)palToMat.size()) { err = "voxel uses material " + std::to_string(m) + " not in the palette"; return VX3H_EPARSE; }
        const int z = (int)((double)i / (double)layer);
        const int y = (int)(((double)i - (double)z * layer) / XV);
        const int x = (int)(i - (size_t)z * layer - (size_t)y * XV);
        o.ijk.push_back(x); o.ijk.push_back(y); o.ijk.push_back(z);
        o.mat.push_back(palToMat[m]);
        voxCell.push_back(i);
    }
    const int NV = (int)o.mat.size();
    o.phase.assign(NV, 0.0);
    if (havePhase) for (int i = 0; i < NV && i < (int)cellPhase.size(); i++) o.phase[i] = cellPhase[i];
    // per voxel cilia vectors (VX_Sim.cpp:102-105), zero where the structure has none
    std::vector<double>* ciliaOut[2] = {&o.baseCilia, &o.shiftCilia};
    for (int c = 0; c < 2; c++) {
        if (cellCilia[c].empty()) continue;
        ciliaOut[c]->assign(3 * (size_t)NV, 0.0);
        for (size_t i = 0; i < ciliaOut[c]->size() && i < cellCilia[c].size(); i++) (*ciliaOut[c])[i] = cellCilia[c][i];
    }

    // boundary conditions -> externals (Import, VX_Sim.cpp:108-141; GetNumTouching, VX_Environment.cpp:215-232)
    if (!bcs.empty()) {
        for (const Region& g : bcs)
            if (g.prim == 3 || g.prim < 0) { err = "mesh (or unreadable) boundary-condition regions are not supported"; return VX3H_EUNSUPPORTED; }
        auto maxOff = [](double lin, double lay, int a, int b, double lim) {
            double m = 0;
            for (int i = 0; i < a; i++)
                for (int j = 0; j < b; j++) { double t = i * lin + j * lay; if (t > m && t < lim) m = t; }
            return m;
        };
        const double W[3] = {latDim * dimAdj[0] * (XV + maxOff(lineOff[0], layerOff[0], YV, ZV, dimAdj[0])),
                             latDim * dimAdj[1] * (YV + maxOff(lineOff[1], layerOff[1], XV, ZV, 1.0)),
                             latDim * dimAdj[2] * ZV};
        const double half[3] = {latEnv[0] / 2.0, latEnv[1] / 2.0, latEnv[2] / 2.0};
        // workspace position of a structure cell, CVX_Object::GetXYZ with lattice offsets (VX_Object.cpp:525-557)
        auto cellXYZ = [&](size_t i, double P[3]) {
            const int iz = (int)((double)i / (double)layer);
            const int iy = (int)(((double)i - (double)iz * layer) / XV);
            const int ix = (int)(i - (size_t)iz * layer - (size_t)iy * XV);
            const double eps = 0.000001;
            double tx = lineOff[0] * iy + layerOff[0] * iz + eps, ty = lineOff[1] * ix + layerOff[1] * iz + eps;
            int bx = (int)(tx / dimAdj[0]), by = (int)(ty / dimAdj[1]);
            tx = lineOff[0] * (iy - by) + layerOff[0] * iz + eps;
            ty = lineOff[1] * (ix - bx) + layerOff[1] * iz + eps;
            bx = (int)(tx / dimAdj[0]);
            by = (int)(ty / dimAdj[1]);
            tx = tx - bx * dimAdj[0];
            ty = ty - by * dimAdj[1];
            P[0] = latDim * (dimAdj[0] * (0.5 + ix) + tx);
            P[1] = latDim * (dimAdj[1] * (0.5 + iy) + ty);
            P[2] = latDim * dimAdj[2] * (0.5 + iz);
        };
        std::vector<int> sizes(bcs.size(), 0);
        for (size_t j = 0; j < bcs.size(); j++)
            for (size_t i = 0; i < cells.size(); i++) {
                if (cells[i] == 0) continue;
                double P[3];
                cellXYZ(i, P);
                if (touching(bcs[j], P, half, W)) sizes[j]++;
            }
        for (int v = 0; v < NV; v++) {
            const double P[3] = {o.ijk[3 * v] * vs + latEnv[0] / 2, o.ijk[3 * v + 1] * vs + latEnv[1] / 2, o.ijk[3 * v + 2] * vs + latEnv[2] / 2};
            bool any = false;
            vx3h_external e;
            memset(&e, 0, sizeof(e));
            e.voxel = v;
            float f[3] = {0, 0, 0}, mo[3] = {0, 0, 0};
            for (size_t j = 0; j < bcs.size(); j++) {
                const Region& g = bcs[j];
                if (!touching(g, P, half, W)) continue;
                any = true;
                for (int k = 0; k < 3; k++) {
                    if (g.dof & (1 << k)) { e.dof_fixed |= 1 << k; if (g.disp[k] != 0.0f) e.translate[k] = g.disp[k]; }
                    if (g.dof & (8 << k)) { e.dof_fixed |= 8 << k; if (g.angDisp[k] != 0.0f) e.rotate[k] = g.angDisp[k]; }
                }
                // external()->addForce(force() + F/n): the current force is added twice
                for (int k = 0; k < 3; k++) {
                    f[k] += (float)((double)f[k] + g.force[k] / sizes[j]);
                    mo[k] += (float)((double)mo[k] + g.torque[k] / sizes[j]);
                }
            }
            if (!any) continue;
            for (int k = 0; k < 3; k++) { e.force[k] = f[k]; e.moment[k] = mo[k]; }
            o.exts.push_back(e);
        }
    }

    // ---- VX3 simulator keys (readVXD, VX3_SimulationManager.cu:323-367)
    if (!parseMathTree(doc, doc.name + ".Simulator.StopCondition.StopConditionFormula", o.stop, err)) return VX3H_EPARSE;
    if (!parseMathTree(doc, doc.name + ".Simulator.FitnessFunction", o.fit, err)) return VX3H_EPARSE;
    const char* ffAxes[3] = {"x_forcefield", "y_forcefield", "z_forcefield"};   // VX3_SimulationManager.cu:349-354
    for (int k = 0; k < 3; k++)
        if (!parseMathTree(doc, doc.name + ".Simulator.ForceField." + ffAxes[k], o.ff[k], err)) return VX3H_EPARSE;
    for (int k = 0; k < 5; k++)   // VX3_SimulationManager.cu:334-343
        if (!parseMathTree(doc, doc.name + ".Simulator.AttachDetach.AttachCondition.Condition_" + std::to_string(k), o.att[k], err))
            return VX3H_EPARSE;
    const std::string S = doc.name + ".Simulator.";
    vx3h_vxa_info& in = o.info;
    memset(&in, 0, sizeof(in));
    in.enable_attach = ptGet<bool>(doc, S + "AttachDetach.EnableAttach", false);
    in.enable_detach = ptGet<bool>(doc, S + "AttachDetach.EnableDetach", false);
    in.enable_cilia = ptGet<int>(doc, S + "EnableCilia", 0);
    in.enable_signals = ptGet<int>(doc, S + "EnableSignals", 0);
    in.secondary_experiment = ptGet<int>(doc, S + "SecondaryExperiment", 0);
    in.record_step_size = ptGet<int>(doc, S + "RecordHistory.RecordStepSize", 0);
    in.max_dist_close_pairs = ptGet<double>(doc, S + "MaxDistInVoxelLengthsToCountAsPair", 0.0);
    in.has_force_field = ptFind(doc, S + "ForceField") != nullptr;
    in.record_voxel = ptGet<int>(doc, S + "RecordHistory.RecordVoxel", 1);
    in.record_link = ptGet<int>(doc, S + "RecordHistory.RecordLink", 0);
    // CVX_Sim::ReadVXA (VXA/VX_Sim.cpp:260-277): read, then ignored by the Voxelyze3 driver
    in.stop_condition_type = ptGet<int>(doc, S + "StopCondition.StopConditionType", 0);
    in.stop_condition_value = ptGet<double>(doc, S + "StopCondition.StopConditionValue", 0.0);
    in.equilibrium_mode = ptGet<int>(doc, S + "EquilibriumMode.EquilibriumModeEnabled", 0) != 0;
    const XNode* rp = ptFind(doc, doc.name + ".RawPrint");
    if (rp) o.rawPrint = rp->data;

    vx3h_sim_desc& d = o.desc;
    memset(&d, 0, sizeof(d));
    d.voxel_size = vs;
    d.n_materials = (int)o.mats.size();
    d.materials = o.mats.data();
    d.n_voxels = NV;
    d.ijk = o.ijk.data();
    d.material = o.mat.data();
    d.phase = o.phase.data();
    d.max_dist_close_pairs = in.max_dist_close_pairs;   // VX3_SimulationManager.cu:358
    d.n_externals = (int)o.exts.size();
    d.externals = o.exts.data();
    d.gravity_mult = (float)(-(double)(float)gravAcc / 9.80665);
    d.floor_enabled = floorEnabled;
    d.temp_enabled = tempEnabled;
    d.vary_temp_enabled = varyTemp;
    d.temp_base = tempBase;
    d.temp_amplitude = tempAmp;
    d.temp_period = tempPeriod;
    d.enable_expansion = ptGet<int>(doc, S + "EnableExpansion", 0);
    d.dt_frac = dtFrac;
    d.dt_fixed = 0;
    d.volume_effects = volumeEffects;
    d.collisions_enabled = ptGet<bool>(doc, S + "AttachDetach.EnableCollision", true);
    d.bounding_radius = ptGet<double>(doc, S + "AttachDetach.boundingRadius", 0.75);
    d.watch_distance = ptGet<double>(doc, S + "AttachDetach.watchDistance", 1.0);
    d.n_stop_tokens = (int)o.stop.size();
    d.stop_tokens = o.stop.data();
    d.n_fitness_tokens = (int)o.fit.size();
    d.fitness_tokens = o.fit.data();
    d.max_steps = 1000000;   // the driver loop's cap (VX3_SimulationManager.cu:62)
    d.save_position_of_all_voxels = ptGet<int>(doc, S + "SavePositionOfAllVoxels", 0);
    d.enable_cilia = in.enable_cilia;
    d.enable_detach = in.enable_detach;
    d.enable_attach = in.enable_attach;
    d.safety_guard = ptGet<int>(doc, S + "AttachDetach.SafetyGuard", 500);   // VX3_SimulationManager.cu:333
    for (int k = 0; k < 5; k++) {
        d.n_attach_tokens[k] = (int)o.att[k].size();
        d.attach_tokens[k] = o.att[k].empty() ? nullptr : o.att[k].data();
    }
    d.secondary_experiment = in.secondary_experiment;
    d.enable_signals = in.enable_signals;
    d.reinit_after = ptGet<double>(doc, S + "ReinitializeInitialPositionAfterThisManySeconds", 0.0);   // VX3_SimulationManager.cu:365
    d.base_cilia_force = o.baseCilia.empty() ? nullptr : o.baseCilia.data();
    d.shift_cilia_force = o.shiftCilia.empty() ? nullptr : o.shiftCilia.data();
    for (int k = 0; k < 3; k++) {
        d.n_force_field_tokens[k] = (int)o.ff[k].size();
        d.force_field_tokens[k] = o.ff[k].empty() ? nullptr : o.ff[k].data();
    }
    return VX3H_OK;
}

std::string fmtDouble(double v) {
    if (std::isnan(v)) return "nan";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char b[40];
    snprintf(b, sizeof b, "%.17g", v);
    return b;
}

std::string xmlEscape(const std::string& s) {
    std::string o;
    for (char c : s) {
        if (c == '<') o += "&lt;";
        else if (c == '>') o += "&gt;";
        else if (c == '&') o += "&amp;";
        else if (c == '"') o += "&quot;";
        else o += c;
    }
    return o;
}

}  // namespace

extern "C" vx3h_status vx3h_vxa_parse(const char* vxa_text, const char* vxd_text, vx3h_vxa** out) {
    if (!vxa_text || !out) { g_err = "null argument"; return VX3H_EINVAL; }
    *out = nullptr;
    XNode doc;
    std::string err;
    if (!parseXml(vxa_text, doc, err)) { g_err = "VXA " + err; return VX3H_EPARSE; }
    if (vxd_text) {
        XNode vxd;
        if (!parseXml(vxd_text, vxd, err)) { g_err = "VXD " + err; return VX3H_EPARSE; }
        mergeVxd(doc, vxd);
    }
    vx3h_vxa* v = new vx3h_vxa();
    vx3h_status st = readVxa(doc, *v, err);
    if (st != VX3H_OK) { g_err = err; delete v; return st; }
    *out = v;
    return VX3H_OK;
}

extern "C" const vx3h_sim_desc* vx3h_vxa_desc(const vx3h_vxa* v) { return v ? &v->desc : nullptr; }

extern "C" vx3h_status vx3h_vxa_get_info(const vx3h_vxa* v, vx3h_vxa_info* info, const char** raw_print) {
    if (!v || !info) { g_err = "null argument"; return VX3H_EINVAL; }
    *info = v->info;
    if (raw_print) *raw_print = v->rawPrint.c_str();
    return VX3H_OK;
}

extern "C" void vx3h_vxa_free(vx3h_vxa* v) { delete v; }

extern "C" vx3h_status vx3h_report_xml(const char* input_dir_name, const vx3h_report_entry* entries, int n, char* buf, size_t cap,
                                       size_t* len) {
    if (n < 0 || (n > 0 && !entries) || !len) { g_err = "bad argument"; return VX3H_EINVAL; }
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    // VX3_SimulationResult::compareFitnessScore: descending, NaN last
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const double fa = entries[a].status->fitness, fb = entries[b].status->fitness;
        if (std::isnan(fb)) return !std::isnan(fa);
        if (std::isnan(fa)) return false;
        return fa > fb;
    });
    std::string x = "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n<report>";
    x += "<inputdir>" + xmlEscape(input_dir_name ? input_dir_name : "") + "</inputdir>";
    if (n > 0) {
        const vx3h_report_entry& b = entries[order[0]];
        x += "<bestfit><filename>" + xmlEscape(b.vxd_name ? b.vxd_name : "") + "</filename><fitness_score>" +
             fmtDouble(b.status->fitness) + "</fitness_score></bestfit>";
    }
    // the property tree keeps one node per simulation name (text before the first '.'):
    // a later entry with the same name overwrites the earlier one's fields in place
    std::vector<std::string> names;
    std::vector<std::string> bodies;
    for (int k = 0; k < n; k++) {
        const vx3h_report_entry& e = entries[order[k]];
        std::string nm = e.vxd_name ? e.vxd_name : "";
        nm = nm.substr(0, nm.find('.'));
        const vx3h_sim_status& s = *e.status;
        std::string body;
        body += "<currentTime>" + fmtDouble(s.current_time) + "</currentTime>";
        body += "<fitness_score>" + fmtDouble(s.fitness) + "</fitness_score>";
        body += "<num_voxel>" + std::to_string(s.num_voxel) + "</num_voxel>";
        body += "<num_measured_voxel>" + std::to_string(s.num_measured_voxel) + "</num_measured_voxel>";
        body += "<voxSize>" + fmtDouble(e.voxel_size) + "</voxSize>";
        body += "<numClosePairs>" + std::to_string(s.num_close_pairs) + "</numClosePairs>";
        body += "<initialCenterOfMass><x>" + fmtDouble(s.initial_com[0]) + "</x><y>" + fmtDouble(s.initial_com[1]) + "</y><z>" +
                fmtDouble(s.initial_com[2]) + "</z></initialCenterOfMass>";
        body += "<currentCenterOfMass><x>" + fmtDouble(s.current_com[0]) + "</x><y>" + fmtDouble(s.current_com[1]) + "</y><z>" +
                fmtDouble(s.current_com[2]) + "</z></currentCenterOfMass>";
        body += "<total_distance_of_all_voxels>" + fmtDouble(s.total_distance_of_all_voxels) + "</total_distance_of_all_voxels>";
        if (e.save_positions && e.init_pos && e.pos && e.mats) {
            auto vec = [&](const double* p) {
                std::string t;
                for (int i = 0; i < s.num_voxel; i++)
                    t += std::to_string(p[3 * i]) + "," + std::to_string(p[3 * i + 1]) + "," + std::to_string(p[3 * i + 2]) + ";";
                return t;
            };
            std::string mats;
            for (int i = 0; i < s.num_voxel; i++) mats += std::to_string(e.mats[i]) + ";";
            body += "<init_pos>" + vec(e.init_pos) + "</init_pos><pos>" + vec(e.pos) + "</pos><mats>" + mats + "</mats>";
        }
        auto it = std::find(names.begin(), names.end(), nm);
        if (it == names.end()) { names.push_back(nm); bodies.push_back(body); }
        else bodies[it - names.begin()] = body;
    }
    x += "<detail>";
    for (size_t i = 0; i < names.size(); i++) x += "<" + names[i] + ">" + bodies[i] + "</" + names[i] + ">";
    x += "</detail></report>";
    *len = x.size() + 1;
    if (buf && cap > 0) {
        const size_t m = std::min(cap - 1, x.size());
        memcpy(buf, x.data(), m);
        buf[m] = 0;
        if (cap < x.size() + 1) { g_err = "buffer too small"; return VX3H_EINVAL; }
    }
    return VX3H_OK;
}
