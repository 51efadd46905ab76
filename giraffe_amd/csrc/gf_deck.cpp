// Reader and writer of GIRAFFE input decks, restricted to the static hot-path
// subset: Nodes, NodeSets, ElementSets, Elements (Beam_1, Shell_1, Pipe_1),
// Materials (Hooke), Sections (Tube, Rectangle), PipeSections, ShellSections (Homogeneous),
// CoordinateSystems, Loads (NodalLoad, NodalFollowerLoad, ShellLoad), Constraints
// (NodalConstraint), Environment (GravityData, OceanData without an active sea
// current), SolutionSteps (Static steps), ConvergenceCriteria.  Output-only blocks (Monitor, PostFiles, SolverOptions,
// ExecutionData, ConfigurationSave, ConcomitantSolution) are skipped; any other top-level block
// changes the physics and is refused with a message naming it.
//
// Grammar as the reference readers take it (token based, "//" starts a comment
// that runs to the end of the line):
//   dispatch            src/IO.cpp:290-657 (top-level keywords)
//   Node / Nodes        src/Node.cpp:152-179
//   NodeSet             src/NodeSet.cpp:30-99 (List or Sequence Initial Increment)
//   ElementSet          src/ElementSet.cpp:30-95
//   Beam_1              src/Beam_1.cpp:370-429 (Mat Sec CS Nodes [PreTension])
//   Shell_1             src/Shell_1.cpp:537-593 (Mat Sec [CS] Nodes n1..n6)
//   Pipe_1              src/Pipe_1.cpp:535-571 (PipeSec CS Nodes)
//   PS (PipeSections)   src/PipeSection.cpp:19-131 (EA EI GJ GA Rho CDt CDn CAt CAn De Di)
//   Hooke               src/Hooke.cpp:16-49 (E Nu Rho)
//   Tube                src/SecTube.cpp:22-44 (De Di), properties :86-106
//   Rectangle           src/SecRectangle.cpp:19-75 (B H [AD ...]), properties :84-97
//   Homogeneous         src/ShellSectionHomogeneous.cpp:17-30 (Thickness)
//   CS                  src/CoordinateSystem.cpp:23-92 (E1 ... E3 ...)
//   NodalLoad           src/NodalLoad.cpp:36-100 (NodeSet CS NTimes + rows t FX..MZ)
//   NodalFollowerLoad   src/NodalFollowerLoad.cpp (same layout)
//   ShellLoad           src/ShellLoad.cpp:28-105 (ElementSet AreaUpdate NTimes + rows t P)
//   NodalConstraint     src/NodalConstraint.cpp:25-140 (NodeSet, DOF BoolTable ...)
//   BoolTable           src/BoolTable.cpp:23-45 (values while the token starts with a digit)
//   GravityData         src/Environment.cpp:355-384 (G gx gy gz BoolTable ...)
//   OceanData           src/Environment.cpp:384-470 (RhoFluid SurfacePosition SeaCurrent N
//                       BoolTable, N x Depth Speed Angle)
//   Static              src/Static.cpp:43-130 (EndTime ... Sample, fixed order)
//   ConvergenceCriteria src/ConvergenceCriteria.cpp:64-165 (12 keys, fixed order)
#include "gf_internal.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <algorithm>
#include <array>
#include <map>
#include <memory>
#include <set>
#include <sstream>

namespace {

const double PI = 3.1415926535897932384626433832795;

// Torsion constant of a B x H rectangle: the series of the reference
// (src/SecRectangle.cpp:91-96), odd n up to 21.
double rect_torsion(double b, double h)
{
    double sum = 0.0;
    for (int n = 1; n < 22; n += 2) sum += std::tanh(n * PI * h / (2.0 * b)) / std::pow((double)n, 5);
    return b * b * b * h / 3.0 * (1.0 - 192.0 * b * sum / (std::pow(PI, 5) * h));
}

// A, I11, I22, It of a section {kind, a, b}: Tube (De, Di), SecTube::PreCalc
// src/SecTube.cpp:86-106; Rectangle (B, H), SecRectangle::PreCalc
// src/SecRectangle.cpp:84-97.
void section_props(const std::array<double, 3> &sec, double &A, double &I11, double &I22, double &It)
{
    const double a = sec[1], b = sec[2];
    if (sec[0] == 0.0) {
        A = PI / 4.0 * (a * a - b * b);
        I11 = I22 = PI / 64.0 * (std::pow(a, 4) - std::pow(b, 4));
        It = PI / 32.0 * (std::pow(a, 4) - std::pow(b, 4));
    } else {
        A = a * b;
        I11 = a * b * b * b / 12.0;
        I22 = b * a * a * a / 12.0;
        It = rect_torsion(a, b);
    }
}

struct Tokens {
    std::vector<std::string> t;
    size_t i = 0;
    bool done() const { return i >= t.size(); }
    const std::string &peek() const { static const std::string e; return i < t.size() ? t[i] : e; }
    std::string next() { return i < t.size() ? t[i++] : std::string(); }
};

bool tokenize(const std::string &path, Tokens &tk)
{
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ls(line);
        std::string w;
        while (ls >> w) {
            if (w.size() >= 2 && w[0] == '/' && w[1] == '/') break;  // comment to end of line
            tk.t.push_back(w);
        }
    }
    return true;
}

const std::set<std::string> &top_keywords()
{
    static const std::set<std::string> k = {
        "SolutionSteps", "Nodes", "SuperNodes", "Points", "Arcs", "Elements", "Particles", "InitialConditions",
        "Materials", "Sections", "PipeSections", "ShellSections", "CoordinateSystems", "RigidBodyData",
        "Environment", "Monitor", "SolverOptions", "AnalyticalSurfaces", "Surfaces", "Splines", "LineRegions",
        "SurfaceRegions", "Contacts", "NodeSets", "SuperNodeSets", "SurfaceSets", "ElementSets", "Loads",
        "Displacements", "Constraints", "SpecialConstraints", "SectionDetails", "AerodynamicData", "BEM",
        "GeneralContactSearch", "ConfigurationSave", "ConcomitantSolution", "ConvergenceCriteria", "PostFiles",
        "ExecutionData", "PSYCoupling", "CADData", "ContactInterfaces", "Boundaries", "Geometries",
        "BodyGeometries", "EOF"};
    return k;
}

struct Elem {
    int type;  // 0 Beam_1, 1 Shell_1, 2 Pipe_1
    int id, mat, sec, cs;
    int nodes[6];
    double T0;
};

struct Load {
    int kind, set, cs, area_update;
    std::vector<double> t, v;  // v: [nt][6]
};

}  // namespace

struct gf_deck {
    // parsed
    std::vector<int> node_ids;
    std::vector<double> node_xyz;
    std::map<int, std::vector<int>> nodesets, elemsets;
    std::vector<Elem> elems;
    std::map<int, std::array<double, 3>> materials;
    std::map<int, std::array<double, 3>> secs;  // kind (0 Tube, 1 Rectangle), De Di | B H
    std::map<int, std::array<double, 7>> pipes;  // PS: EA EI GJ GA Rho De Di
    int ocean_on = 0;
    double rho_fluid = 0.0, surface[3] = {0, 0, 0};
    std::map<int, double> shell_t;
    std::map<int, std::array<double, 6>> css;
    std::vector<Load> loads;
    std::vector<std::pair<int, std::array<int, 6>>> constraints;  // values in every step
    std::vector<std::pair<int, std::array<std::vector<int>, 6>>> ctables;  // BoolTables as read
    double G[3] = {0, 0, 0};
    int g_on = 0;
    std::vector<int> g_table;  // gravity BoolTable
    bool current_moving = false;
    std::vector<int> current_table;  // sea-current BoolTable (when its speed is non-zero)
    gf_static_params p{};     // ConvergenceCriteria (and the first step after resolve)
    std::vector<gf_static_params> steps;
    std::vector<int> g_mode;  // GF_GRAVITY_* per step
    bool have_static = false;
    // resolved descriptor arrays
    std::vector<unsigned char> fix;
    std::vector<int> beam_conn, shell_conn, load_kind, load_set_off, load_set_items, load_flags, load_tab_off;
    std::vector<int> beam_number, shell_number, beam_kind;
    std::vector<double> beam_buoy;
    std::vector<double> beam_cs, beam_prop, shell_prop, load_cs, tab_time, tab_val;
    gf_model_desc desc{};
};

namespace {

int fail(const std::string &msg)
{
    gf_set_error("deck: " + msg);
    return GF_E_ARG;
}

bool expect(Tokens &tk, const char *kw) { return tk.next() == kw; }

// BoolTable::GetAt src/BoolTable.cpp:69-82: entry i, the last entry past the
// end, false before the first (and for an empty table).
bool bool_at(const std::vector<int> &t, int i)
{
    if (i < 0 || t.empty()) return false;
    return t[std::min((size_t)i, t.size() - 1)] != 0;
}

// src/BoolTable.cpp:23-45: values while the next token starts with a digit
std::vector<int> read_booltable(Tokens &tk)
{
    std::vector<int> v;
    while (!tk.done() && std::isdigit((unsigned char)tk.peek()[0])) v.push_back(std::atoi(tk.next().c_str()) == 1);
    return v;
}

int read_set(Tokens &tk, const char *head, const char *items, std::map<int, std::vector<int>> &out)
{
    if (!expect(tk, head)) return fail(std::string("expected ") + head);
    int id = std::atoi(tk.next().c_str());
    if (!expect(tk, items)) return fail(std::string(head) + ": expected " + items);
    int n = std::atoi(tk.next().c_str());
    std::string how = tk.next();
    std::vector<int> v;
    if (how == "List") {
        for (int i = 0; i < n; i++) v.push_back(std::atoi(tk.next().c_str()));
    } else if (how == "Sequence") {
        if (!expect(tk, "Initial")) return fail(std::string(head) + ": expected Initial");
        int a = std::atoi(tk.next().c_str());
        if (!expect(tk, "Increment")) return fail(std::string(head) + ": expected Increment");
        int inc = std::atoi(tk.next().c_str());
        for (int i = 0; i < n; i++) v.push_back(a + i * inc);
    } else {
        return fail(std::string(head) + ": expected List or Sequence");
    }
    out[id] = v;
    return GF_OK;
}

int parse(Tokens &tk, gf_deck &d)
{
    const std::set<std::string> &top = top_keywords();
    // output only: ConcomitantSolution runs a Modal analysis at sampled
    // converged configurations and writes its eigenvalues to a file
    // (ConcomitantSolution::UpdateConcomitantSolution src/ConcomitantSolution.cpp:106-124);
    // it does not change the static solution, and its file is not written here
    const std::set<std::string> ignored = {"Monitor", "SolverOptions", "PostFiles", "ExecutionData",
                                           "ConfigurationSave", "ConcomitantSolution"};
    while (!tk.done()) {
        std::string kw = tk.next();
        if (kw == "EOF") break;
        if (!top.count(kw)) return fail("unexpected token '" + kw + "'");
        if (ignored.count(kw)) {
            while (!tk.done() && !top.count(tk.peek())) tk.next();
            continue;
        }
        if (kw == "Nodes") {
            int n = std::atoi(tk.next().c_str());
            for (int i = 0; i < n; i++) {
                if (!expect(tk, "Node")) return fail("Nodes: expected Node");
                d.node_ids.push_back(std::atoi(tk.next().c_str()));
                for (int k = 0; k < 3; k++) d.node_xyz.push_back(std::atof(tk.next().c_str()));
            }
        } else if (kw == "NodeSets") {
            int n = std::atoi(tk.next().c_str());
            for (int i = 0; i < n; i++) GF_TRY(read_set(tk, "NodeSet", "Nodes", d.nodesets));
        } else if (kw == "ElementSets") {
            int n = std::atoi(tk.next().c_str());
            for (int i = 0; i < n; i++) GF_TRY(read_set(tk, "ElementSet", "Elements", d.elemsets));
        } else if (kw == "Elements") {
            int n = std::atoi(tk.next().c_str());
            for (int i = 0; i < n; i++) {
                std::string type = tk.next();
                Elem e{};
                if (type == "Beam_1") e.type = 0;
                else if (type == "Shell_1") e.type = 1;
                else if (type == "Pipe_1") e.type = 2;
                else return fail("element type " + type + " is not on the static hot path");
                e.id = std::atoi(tk.next().c_str());
                if (e.type == 2) {
                    // Pipe_1::Read src/Pipe_1.cpp:535-571: PipeSec s CS c Nodes a b c
                    if (!expect(tk, "PipeSec")) return fail("Pipe_1: expected PipeSec");
                    e.sec = std::atoi(tk.next().c_str());
                    if (!expect(tk, "CS")) return fail("Pipe_1: expected CS");
                    e.cs = std::atoi(tk.next().c_str());
                    if (!expect(tk, "Nodes")) return fail("Pipe_1: expected Nodes");
                    for (int a = 0; a < 3; a++) e.nodes[a] = std::atoi(tk.next().c_str());
                    d.elems.push_back(e);
                    continue;
                }
                if (!expect(tk, "Mat")) return fail(type + ": expected Mat");
                e.mat = std::atoi(tk.next().c_str());
                if (!expect(tk, "Sec")) return fail(type + ": expected Sec");
                e.sec = std::atoi(tk.next().c_str());
                e.cs = 0;
                if (e.type == 0) {
                    if (!expect(tk, "CS")) return fail("Beam_1: expected CS");
                    e.cs = std::atoi(tk.next().c_str());
                } else if (tk.peek() == "CS") {
                    tk.next();
                    e.cs = std::atoi(tk.next().c_str());
                }
                if (!expect(tk, "Nodes")) return fail(type + ": expected Nodes");
                int nn = e.type == 0 ? 3 : 6;
                for (int a = 0; a < nn; a++) e.nodes[a] = std::atoi(tk.next().c_str());
                if (e.type == 0 && tk.peek() == "PreTension") {
                    tk.next();
                    e.T0 = std::atof(tk.next().c_str());
                }
                d.elems.push_back(e);
            }
        } else if (kw == "Materials") {
            int n = std::atoi(tk.next().c_str());
            for (int i = 0; i < n; i++) {
                std::string type = tk.next();
                if (type != "Hooke") return fail("material " + type + " is not on the static hot path");
                int id = std::atoi(tk.next().c_str());
                std::array<double, 3> m{};
                if (!expect(tk, "E")) return fail("Hooke: expected E");
                m[0] = std::atof(tk.next().c_str());
                if (!expect(tk, "Nu")) return fail("Hooke: expected Nu");
                m[1] = std::atof(tk.next().c_str());
                if (!expect(tk, "Rho")) return fail("Hooke: expected Rho");
                m[2] = std::atof(tk.next().c_str());
                d.materials[id] = m;
            }
        } else if (kw == "Sections") {
            int n = std::atoi(tk.next().c_str());
            for (int i = 0; i < n; i++) {
                std::string type = tk.next();
                if (type != "Tube" && type != "Rectangle")
                    return fail("section " + type + " is not on the static hot path (Tube and Rectangle only)");
                const bool tube = type == "Tube";
                int id = std::atoi(tk.next().c_str());
                if (!expect(tk, tube ? "De" : "B")) return fail(type + (tube ? ": expected De" : ": expected B"));
                double a = std::atof(tk.next().c_str());
                if (!expect(tk, tube ? "Di" : "H")) return fail(type + (tube ? ": expected Di" : ": expected H"));
                double b = std::atof(tk.next().c_str());
                if (tk.peek() == "AD") return fail(type + ": aerodynamic data is not on the static hot path");
                d.secs[id] = {tube ? 0.0 : 1.0, a, b};
            }
        } else if (kw == "PipeSections") {
            // PipeSection::Read src/PipeSection.cpp:19-131
            static const char *keys[11] = {"EA", "EI", "GJ", "GA", "Rho", "CDt", "CDn", "CAt", "CAn", "De", "Di"};
            int n = std::atoi(tk.next().c_str());
            for (int i = 0; i < n; i++) {
                if (!expect(tk, "PS")) return fail("PipeSections: expected PS");
                int id = std::atoi(tk.next().c_str());
                double v[11];
                for (int k = 0; k < 11; k++) {
                    if (!expect(tk, keys[k])) return fail(std::string("PS: expected ") + keys[k]);
                    v[k] = std::atof(tk.next().c_str());
                }
                d.pipes[id] = {v[0], v[1], v[2], v[3], v[4], v[9], v[10]};
            }
        } else if (kw == "ShellSections") {
            int n = std::atoi(tk.next().c_str());
            for (int i = 0; i < n; i++) {
                std::string type = tk.next();
                if (type != "Homogeneous") return fail("shell section " + type + " is not on the static hot path");
                int id = std::atoi(tk.next().c_str());
                if (!expect(tk, "Thickness")) return fail("Homogeneous: expected Thickness");
                d.shell_t[id] = std::atof(tk.next().c_str());
            }
        } else if (kw == "CoordinateSystems") {
            int n = std::atoi(tk.next().c_str());
            for (int i = 0; i < n; i++) {
                if (!expect(tk, "CS")) return fail("CoordinateSystems: expected CS");
                int id = std::atoi(tk.next().c_str());
                std::array<double, 6> c{};
                if (!expect(tk, "E1")) return fail("CS: expected E1");
                for (int k = 0; k < 3; k++) c[k] = std::atof(tk.next().c_str());
                if (!expect(tk, "E3")) return fail("CS: expected E3");
                for (int k = 0; k < 3; k++) c[3 + k] = std::atof(tk.next().c_str());
                d.css[id] = c;
            }
        } else if (kw == "Loads") {
            int n = std::atoi(tk.next().c_str());
            for (int i = 0; i < n; i++) {
                std::string type = tk.next();
                Load L{};
                tk.next();  // load id
                if (type == "NodalLoad" || type == "NodalFollowerLoad") {
                    L.kind = type == "NodalLoad" ? GF_LOAD_NODAL : GF_LOAD_NODAL_FOLLOWER;
                    if (!expect(tk, "NodeSet")) return fail(type + ": expected NodeSet");
                    L.set = std::atoi(tk.next().c_str());
                    if (!expect(tk, "CS")) return fail(type + ": expected CS");
                    L.cs = std::atoi(tk.next().c_str());
                } else if (type == "ShellLoad") {
                    L.kind = GF_LOAD_SHELL_PRESSURE;
                    if (!expect(tk, "ElementSet")) return fail("ShellLoad: expected ElementSet");
                    L.set = std::atoi(tk.next().c_str());
                    if (!expect(tk, "AreaUpdate")) return fail("ShellLoad: expected AreaUpdate");
                    L.area_update = std::atoi(tk.next().c_str());
                } else {
                    return fail("load " + type + " is not on the static hot path");
                }
                std::string nt_kw = tk.next();
                if (nt_kw == "MathCode") return fail(type + ": MathCode loads are not supported");
                if (nt_kw != "NTimes") return fail(type + ": expected NTimes");
                int nt = std::atoi(tk.next().c_str());
                for (int r = 0; r < nt; r++) {
                    L.t.push_back(std::atof(tk.next().c_str()));
                    int nv = L.kind == GF_LOAD_SHELL_PRESSURE ? 1 : 6;
                    for (int k = 0; k < 6; k++) L.v.push_back(k < nv ? std::atof(tk.next().c_str()) : 0.0);
                }
                d.loads.push_back(L);
            }
        } else if (kw == "Constraints") {
            int n = std::atoi(tk.next().c_str());
            for (int i = 0; i < n; i++) {
                std::string type = tk.next();
                if (type != "NodalConstraint") return fail("constraint " + type + " is not on the static hot path");
                tk.next();  // id
                if (!expect(tk, "NodeSet")) return fail("NodalConstraint: expected NodeSet");
                int set = std::atoi(tk.next().c_str());
                std::array<std::vector<int>, 6> f{};
                static const char *dofs[6] = {"UX", "UY", "UZ", "ROTX", "ROTY", "ROTZ"};
                for (;;) {
                    int k = -1;
                    for (int q = 0; q < 6; q++)
                        if (tk.peek() == dofs[q]) k = q;
                    if (k < 0) break;
                    tk.next();
                    if (!expect(tk, "BoolTable")) return fail("NodalConstraint: expected BoolTable");
                    f[k] = read_booltable(tk);
                }
                d.ctables.push_back({set, f});
            }
        } else if (kw == "Environment") {
            while (!tk.done() && !top.count(tk.peek())) {
                std::string e = tk.next();
                if (e == "OceanData") {
                    // Environment::Read src/Environment.cpp:384-470
                    if (!expect(tk, "RhoFluid")) return fail("OceanData: expected RhoFluid");
                    d.rho_fluid = std::atof(tk.next().c_str());
                    if (!expect(tk, "SurfacePosition")) return fail("OceanData: expected SurfacePosition");
                    for (int k = 0; k < 3; k++) d.surface[k] = std::atof(tk.next().c_str());
                    if (!expect(tk, "SeaCurrent")) return fail("OceanData: expected SeaCurrent");
                    if (!expect(tk, "N")) return fail("SeaCurrent: expected N");
                    const int np = std::atoi(tk.next().c_str());
                    if (!expect(tk, "BoolTable")) return fail("SeaCurrent: expected BoolTable");
                    std::vector<int> b = read_booltable(tk);
                    bool moving = false;
                    for (int i = 0; i < np; i++) {
                        static const char *pk[3] = {"Depth", "Speed", "Angle"};
                        double v[3];
                        for (int k = 0; k < 3; k++) {
                            if (!expect(tk, pk[k])) return fail(std::string("SeaCurrent: expected ") + pk[k]);
                            v[k] = std::atof(tk.next().c_str());
                        }
                        moving = moving || v[1] != 0.0;
                    }
                    d.current_moving = moving;
                    d.current_table = b;
                    d.ocean_on = 1;
                    continue;
                }
                if (e != "GravityData") return fail("environment data " + e + " is not on the static hot path");
                if (!expect(tk, "G")) return fail("GravityData: expected G");
                for (int k = 0; k < 3; k++) d.G[k] = std::atof(tk.next().c_str());
                if (!expect(tk, "BoolTable")) return fail("GravityData: expected BoolTable");
                d.g_table = read_booltable(tk);
            }
        } else if (kw == "SolutionSteps") {
            // Static::Read src/Static.cpp:43-130, one per step; start times follow
            // Database::EvaluateStartEndTimes src/Database.cpp:760-770
            int n = std::atoi(tk.next().c_str());
            if (n < 1) return fail("SolutionSteps: no solution step");
            double start = 0.0;
            for (int i = 0; i < n; i++) {
                std::string type = tk.next();
                if (type != "Static") return fail("solution step " + type + " is not on the static hot path");
                tk.next();  // id
                static const char *keys[9] = {"EndTime", "TimeStep", "MaxTimeStep", "MinTimeStep", "MaxIt",
                                              "MinIt", "ConvIncrease", "IncFactor", "Sample"};
                double v[9];
                for (int k = 0; k < 9; k++) {
                    if (!expect(tk, keys[k])) return fail(std::string("Static: expected ") + keys[k]);
                    v[k] = std::atof(tk.next().c_str());
                }
                gf_static_params q{};
                q.start_time = start;
                q.end_time = v[0];
                q.time_step = v[1];
                q.max_time_step = v[2];
                q.min_time_step = v[3];
                q.max_it = (int)v[4];
                q.min_it = (int)v[5];
                q.conv_increase = (int)v[6];
                q.inc_factor = v[7];
                q.sample = (int)v[8];
                d.steps.push_back(q);
                start = q.end_time;
            }
            d.have_static = true;
        } else if (kw == "ConvergenceCriteria") {
            static const char *keys[12] = {"ForceTolerance", "MomentTolerance", "ForceMinimumReference",
                                           "MomentMinimumReference", "ConstraintMinimumReference",
                                           "DisplacementTolerance", "RotationTolerance", "LagrangeTolerance",
                                           "DisplacementMinimumReference", "RotationMinimumReference",
                                           "LagrangeMinimumReference", "DivergenceReference"};
            double *dst[12] = {&d.p.force_tol, &d.p.moment_tol, &d.p.force_min, &d.p.moment_min,
                               &d.p.constraint_min, &d.p.disp_tol, &d.p.rot_tol, &d.p.lag_tol, &d.p.disp_min,
                               &d.p.rot_min, &d.p.lag_min, &d.p.divergence_ref};
            for (int k = 0; k < 12; k++) {
                if (!expect(tk, keys[k])) return fail(std::string("ConvergenceCriteria: expected ") + keys[k]);
                *dst[k] = std::atof(tk.next().c_str());
            }
        } else {
            return fail("block " + kw + " is not on the static hot path");
        }
    }
    if (!d.have_static) return fail("no Static solution step");
    // per step: the criteria (one ConvergenceCriteria for all steps), the gravity
    // factor mode from the table's entries for this and the previous step
    // (BoolTable::GetAt src/BoolTable.cpp:69-82: past the end the last entry,
    // before the first false), the constraints, which must not change
    const int n = (int)d.steps.size();
    d.g_on = 0;
    for (int i = 0; i < n; i++) {
        gf_static_params &q = d.steps[i];
        const double t0 = q.start_time, t1 = q.end_time, dt = q.time_step, dtmax = q.max_time_step,
                     dtmin = q.min_time_step, fi = q.inc_factor;
        const int mi = q.max_it, mn = q.min_it, ci = q.conv_increase, sa = q.sample;
        q = d.p;
        q.start_time = t0;
        q.end_time = t1;
        q.time_step = dt;
        q.max_time_step = dtmax;
        q.min_time_step = dtmin;
        q.max_it = mi;
        q.min_it = mn;
        q.conv_increase = ci;
        q.inc_factor = fi;
        q.sample = sa;
        const bool now = bool_at(d.g_table, i), before = bool_at(d.g_table, i - 1);
        d.g_mode.push_back(now ? (before ? GF_GRAVITY_ON : GF_GRAVITY_RAMP_UP)
                               : (before ? GF_GRAVITY_RAMP_DOWN : GF_GRAVITY_OFF));
        if (now) d.g_on = 1;
        if (d.current_moving && bool_at(d.current_table, i))
            return fail("SeaCurrent: Morison loads of a sea current are not on the static hot path");
    }
    for (auto &c : d.ctables) {
        std::array<int, 6> f{};
        for (int k = 0; k < 6; k++) {
            f[k] = bool_at(c.second[k], 0) ? 1 : 0;
            for (int i = 1; i < n; i++)
                if (bool_at(c.second[k], i) != (f[k] != 0))
                    return fail("constraints that change between solution steps are not supported");
        }
        d.constraints.push_back({c.first, f});
    }
    d.p = d.steps[0];
    return GF_OK;
}

// CoordinateSystem::Read: rows E1, E2 = E3 x E1, E3, each normalised.
void cs_rows(const std::array<double, 6> &c, double *out)
{
    double e1[3] = {c[0], c[1], c[2]}, e3[3] = {c[3], c[4], c[5]};
    double e2[3] = {e3[1] * e1[2] - e3[2] * e1[1], e3[2] * e1[0] - e3[0] * e1[2], e3[0] * e1[1] - e3[1] * e1[0]};
    double *rows[3] = {e1, e2, e3};
    for (int r = 0; r < 3; r++) {
        double n = std::sqrt(rows[r][0] * rows[r][0] + rows[r][1] * rows[r][1] + rows[r][2] * rows[r][2]);
        for (int k = 0; k < 3; k++) out[3 * r + k] = n != 1.0 ? rows[r][k] / n : rows[r][k];
    }
}

int resolve(gf_deck &d)
{
    const int nn = (int)d.node_ids.size();
    std::map<int, int> nidx;
    for (int i = 0; i < nn; i++) nidx[d.node_ids[i]] = i;
    auto node = [&](int id, int &out) -> int {
        auto it = nidx.find(id);
        if (it == nidx.end()) return fail("node " + std::to_string(id) + " is not defined");
        out = it->second;
        return GF_OK;
    };
    std::map<int, int> shell_of;  // element id -> shell index
    for (const Elem &e : d.elems) {
        if (e.type == 2) {
            // Pipe_1::PreCalc src/Pipe_1.cpp:1106-1160: D and Rho from the pipe section,
            // Aext = pi De^2 / 4 for the buoyancy below the surface
            auto pi = d.pipes.find(e.sec);
            if (pi == d.pipes.end()) return fail("pipe section " + std::to_string(e.sec) + " is not defined");
            auto ci = d.css.find(e.cs);
            if (ci == d.css.end()) return fail("coordinate system " + std::to_string(e.cs) + " is not defined");
            for (int a = 0; a < 3; a++) {
                int k;
                GF_TRY(node(e.nodes[a], k));
                d.beam_conn.push_back(k);
            }
            d.beam_number.push_back(e.id);
            double cs[9];
            cs_rows(ci->second, cs);
            d.beam_cs.insert(d.beam_cs.end(), cs, cs + 9);
            const std::array<double, 7> &ps = pi->second;
            double prop[GF_BEAM_NPROP] = {ps[0], ps[1], ps[2], ps[3], ps[4], 0.0, 0.0, 0.0, 0.0};
            d.beam_prop.insert(d.beam_prop.end(), prop, prop + GF_BEAM_NPROP);
            d.beam_kind.push_back(GF_BEAM_KIND_PIPE1);
            d.beam_buoy.push_back(d.ocean_on ? d.rho_fluid * (PI * ps[5] * ps[5] / 4.0) : 0.0);
            continue;
        }
        auto mi = d.materials.find(e.mat);
        if (mi == d.materials.end()) return fail("material " + std::to_string(e.mat) + " is not defined");
        if (e.type == 0) {
            d.beam_kind.push_back(GF_BEAM_KIND_BEAM1);
            d.beam_buoy.push_back(0.0);
            auto si = d.secs.find(e.sec);
            if (si == d.secs.end()) return fail("section " + std::to_string(e.sec) + " is not defined");
            auto ci = d.css.find(e.cs);
            if (ci == d.css.end()) return fail("coordinate system " + std::to_string(e.cs) + " is not defined");
            for (int a = 0; a < 3; a++) {
                int k;
                GF_TRY(node(e.nodes[a], k));
                d.beam_conn.push_back(k);
            }
            d.beam_number.push_back(e.id);
            double cs[9];
            cs_rows(ci->second, cs);
            d.beam_cs.insert(d.beam_cs.end(), cs, cs + 9);
            double A, I11, I22, It;
            section_props(si->second, A, I11, I22, It);
            double prop[GF_BEAM_NPROP] = {mi->second[0], mi->second[1], mi->second[2], A, I11, I22, 0.0, It, e.T0};
            d.beam_prop.insert(d.beam_prop.end(), prop, prop + GF_BEAM_NPROP);
        } else {
            auto ti = d.shell_t.find(e.sec);
            if (ti == d.shell_t.end()) return fail("shell section " + std::to_string(e.sec) + " is not defined");
            shell_of[e.id] = (int)d.shell_conn.size() / 6;
            for (int a = 0; a < 6; a++) {
                int k;
                GF_TRY(node(e.nodes[a], k));
                d.shell_conn.push_back(k);
            }
            d.shell_number.push_back(e.id);
            double prop[GF_SHELL_NPROP] = {mi->second[0], mi->second[1], mi->second[2], ti->second};
            d.shell_prop.insert(d.shell_prop.end(), prop, prop + GF_SHELL_NPROP);
        }
    }
    d.fix.assign(6 * (size_t)nn, 0);
    for (auto &c : d.constraints) {
        auto si = d.nodesets.find(c.first);
        if (si == d.nodesets.end()) return fail("node set " + std::to_string(c.first) + " is not defined");
        for (int id : si->second) {
            int k;
            GF_TRY(node(id, k));
            for (int q = 0; q < 6; q++)
                if (c.second[q]) d.fix[6 * (size_t)k + q] = 1;
        }
    }
    d.load_set_off.push_back(0);
    d.load_tab_off.push_back(0);
    for (const Load &L : d.loads) {
        d.load_kind.push_back(L.kind);
        d.load_flags.push_back(L.area_update);
        double cs[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        if (L.kind == GF_LOAD_SHELL_PRESSURE) {
            auto si = d.elemsets.find(L.set);
            if (si == d.elemsets.end()) return fail("element set " + std::to_string(L.set) + " is not defined");
            for (int id : si->second) {
                auto it = shell_of.find(id);
                if (it == shell_of.end()) return fail("ShellLoad on element " + std::to_string(id) + ", not a Shell_1");
                d.load_set_items.push_back(it->second);
            }
        } else {
            auto si = d.nodesets.find(L.set);
            if (si == d.nodesets.end()) return fail("node set " + std::to_string(L.set) + " is not defined");
            for (int id : si->second) {
                int k;
                GF_TRY(node(id, k));
                d.load_set_items.push_back(k);
            }
            auto ci = d.css.find(L.cs);
            if (ci == d.css.end()) return fail("coordinate system " + std::to_string(L.cs) + " is not defined");
            cs_rows(ci->second, cs);
        }
        d.load_cs.insert(d.load_cs.end(), cs, cs + 9);
        d.load_set_off.push_back((int)d.load_set_items.size());
        d.tab_time.insert(d.tab_time.end(), L.t.begin(), L.t.end());
        d.tab_val.insert(d.tab_val.end(), L.v.begin(), L.v.end());
        d.load_tab_off.push_back((int)d.tab_time.size());
    }
    gf_model_desc &g = d.desc;
    g = gf_model_desc{};
    g.n_nodes = nn;
    g.node_xyz = d.node_xyz.data();
    g.node_fix = d.fix.data();
    g.n_beams = (int)d.beam_conn.size() / 3;
    g.beam_conn = d.beam_conn.data();
    g.beam_cs = d.beam_cs.data();
    g.beam_prop = d.beam_prop.data();
    g.n_shells = (int)d.shell_conn.size() / 6;
    g.shell_conn = d.shell_conn.data();
    g.shell_prop = d.shell_prop.data();
    g.n_loads = (int)d.loads.size();
    g.load_kind = d.load_kind.data();
    g.load_set_off = d.load_set_off.data();
    g.load_set_items = d.load_set_items.data();
    g.load_cs = d.load_cs.data();
    g.load_flags = d.load_flags.data();
    g.load_tab_off = d.load_tab_off.data();
    g.tab_time = d.tab_time.data();
    g.tab_val = d.tab_val.data();
    for (int k = 0; k < 3; k++) g.gravity[k] = d.G[k];
    {
        // element numbers for the results writer, when they number the elements 1..N
        const size_t ne = d.beam_number.size() + d.shell_number.size();
        std::vector<char> seen(ne + 1, 0);
        bool perm = true;
        for (const std::vector<int> *v : {&d.beam_number, &d.shell_number})
            for (int id : *v) {
                if (id < 1 || (size_t)id > ne || seen[id]) perm = false;
                else seen[id] = 1;
            }
        g.beam_number = perm ? d.beam_number.data() : nullptr;
        g.shell_number = perm ? d.shell_number.data() : nullptr;
    }
    g.beam_kind = d.beam_kind.data();
    g.beam_buoy = d.beam_buoy.data();
    g.ocean_on = d.ocean_on;
    for (int k = 0; k < 3; k++) g.ocean_surface[k] = d.surface[k];
    g.ocean_rho_fluid = d.rho_fluid;
    g.gravity_on = d.g_on;
    return GF_OK;
}

std::string num(double v)
{
    char b[40];
    std::snprintf(b, sizeof(b), "%.17g", v);
    return b;
}

}  // namespace

extern "C" int gf_deck_read(const char *path, gf_deck **out)
{
    if (!path || !out) return GF_E_ARG;
    auto d = std::make_unique<gf_deck>();
    // ConvergenceCriteria defaults (src/ConvergenceCriteria.cpp:15-40)
    d->p.force_tol = d->p.moment_tol = 1e-4;
    d->p.force_min = d->p.moment_min = 1e-5;
    d->p.constraint_min = 1e-7;
    d->p.disp_tol = d->p.rot_tol = d->p.lag_tol = 1e-4;
    d->p.disp_min = d->p.rot_min = d->p.lag_min = 1e-6;
    d->p.divergence_ref = 1e15;
    Tokens tk;
    if (!tokenize(path, tk)) return fail(std::string("cannot open ") + path);
    GF_TRY(parse(tk, *d));
    GF_TRY(resolve(*d));
    *out = d.release();
    return GF_OK;
}

extern "C" int gf_deck_desc(const gf_deck *d, gf_model_desc *desc, gf_static_params *p)
{
    if (!d) return GF_E_ARG;
    if (desc) *desc = d->desc;
    if (p) *p = d->p;
    return GF_OK;
}

extern "C" int gf_deck_steps(const gf_deck *d, int *n)
{
    if (!d || !n) return GF_E_ARG;
    *n = (int)d->steps.size();
    return GF_OK;
}

extern "C" int gf_deck_step(const gf_deck *d, int i, gf_static_params *p, int *gravity_mode)
{
    if (!d || i < 0 || i >= (int)d->steps.size()) return GF_E_ARG;
    if (p) *p = d->steps[i];
    if (gravity_mode) *gravity_mode = d->g_mode[i];
    return GF_OK;
}

extern "C" int gf_deck_destroy(gf_deck *d)
{
    delete d;
    return GF_OK;
}

// Writes the description as a deck of the same subset: one Hooke / Tube / Rectangle or
// Homogeneous / CS per distinct combination, node sets per load and per
// constraint pattern, element sets per shell load.
extern "C" int gf_deck_write(const char *path, const gf_model_desc *g, const gf_static_params *p)
{
    if (!path || !g || !p) return GF_E_ARG;
    std::ostringstream o;
    o << "//GIRAFFE deck: static hot-path subset\n";
    const int nn = g->n_nodes, nb = g->n_beams, ns = g->n_shells, nl = g->n_loads;
    o << "Nodes\t" << nn << "\n";
    for (int i = 0; i < nn; i++)
        o << "Node\t" << i + 1 << "\t" << num(g->node_xyz[3 * i]) << "\t" << num(g->node_xyz[3 * i + 1]) << "\t"
          << num(g->node_xyz[3 * i + 2]) << "\n";
    // materials, sections, coordinate systems (deduplicated)
    std::map<std::array<double, 3>, int> mats;
    std::map<std::array<double, 3>, int> tubes;  // kind, De Di | B H
    std::map<double, int> thick;
    std::map<std::array<double, 6>, int> css;
    auto mat_id = [&](double E, double nu, double rho) {
        std::array<double, 3> k{E, nu, rho};
        auto it = mats.find(k);
        if (it != mats.end()) return it->second;
        int id = (int)mats.size() + 1;
        mats[k] = id;
        return id;
    };
    auto cs_id = [&](const double *rows) {
        std::array<double, 6> k{rows[0], rows[1], rows[2], rows[6], rows[7], rows[8]};
        auto it = css.find(k);
        if (it != css.end()) return it->second;
        int id = (int)css.size() + 1;
        css[k] = id;
        return id;
    };
    // element lines in element-number order (numbers from the description or 1..nb, nb+1..)
    auto beam_num = [&](int e) { return g->beam_number ? g->beam_number[e] : e + 1; };
    auto shell_num = [&](int e) { return g->shell_number ? g->shell_number[e] : nb + e + 1; };
    std::vector<std::string> lines((size_t)nb + ns);
    std::map<std::array<double, 6>, int> pipes;  // EA EI GJ GA Rho De -> PS id
    for (int e = 0; e < nb; e++) {
        std::ostringstream el;
        const double *pr = g->beam_prop + GF_BEAM_NPROP * (size_t)e;
        if (g->beam_kind && g->beam_kind[e] == GF_BEAM_KIND_PIPE1) {
            // De back from the buoyant mass per length rho_f pi De^2 / 4 (0 without OceanData)
            const double buoy = g->beam_buoy ? g->beam_buoy[e] : 0.0;
            const double De = g->ocean_on && g->ocean_rho_fluid > 0 ? std::sqrt(4.0 * buoy / (PI * g->ocean_rho_fluid)) : 0.0;
            std::array<double, 6> key{pr[0], pr[1], pr[2], pr[3], pr[4], De};
            int sec;
            auto it = pipes.find(key);
            if (it != pipes.end()) sec = it->second;
            else { sec = (int)pipes.size() + 1; pipes[key] = sec; }
            el << "Pipe_1\t" << beam_num(e) << "\tPipeSec\t" << sec << "\tCS\t" << cs_id(g->beam_cs + 9 * (size_t)e)
               << "\tNodes";
            for (int a = 0; a < 3; a++) el << "\t" << g->beam_conn[3 * e + a] + 1;
            el << "\n";
            const int k = beam_num(e) - 1;
            if (k < 0 || k >= nb + ns || !lines[k].empty()) return fail("element numbers are not a permutation of 1..N");
            lines[k] = el.str();
            continue;
        }
        double A = pr[3], I = pr[4];
        std::array<double, 3> tk;
        if (pr[6] == 0.0 && pr[4] == pr[5] && std::fabs(pr[7] - 2.0 * I) <= 1e-12 * I) {
            double s = 16.0 * I / A, dd = 4.0 * A / PI;  // De^2 + Di^2, De^2 - Di^2
            tk = {0.0, std::sqrt(0.5 * (s + dd)), std::sqrt(std::fmax(0.0, 0.5 * (s - dd)))};
        } else {
            // a B x H rectangle: I11 / I22 = (H / B)^2, A = B H
            const double r = std::sqrt(pr[4] / pr[5]), B = std::sqrt(A / r);
            tk = {1.0, B, A / B};
            double A2, I11, I22, It;
            section_props(tk, A2, I11, I22, It);
            if (pr[6] != 0.0 || std::fabs(A2 - A) > 1e-12 * A || std::fabs(I11 - pr[4]) > 1e-12 * pr[4] ||
                std::fabs(I22 - pr[5]) > 1e-12 * pr[5] || std::fabs(It - pr[7]) > 1e-9 * pr[7])
                return fail("beam " + std::to_string(e) + " is neither a tube nor a rectangle section");
        }
        int sec;
        auto it = tubes.find(tk);
        if (it != tubes.end()) sec = it->second;
        else { sec = (int)tubes.size() + 1; tubes[tk] = sec; }
        el << "Beam_1\t" << beam_num(e) << "\tMat\t" << mat_id(pr[0], pr[1], pr[2]) << "\tSec\t" << sec << "\tCS\t"
           << cs_id(g->beam_cs + 9 * (size_t)e) << "\tNodes";
        for (int a = 0; a < 3; a++) el << "\t" << g->beam_conn[3 * e + a] + 1;
        if (pr[8] != 0.0) el << "\tPreTension\t" << num(pr[8]);
        el << "\n";
        const int k = beam_num(e) - 1;
        if (k < 0 || k >= nb + ns || !lines[k].empty()) return fail("element numbers are not a permutation of 1..N");
        lines[k] = el.str();
    }
    for (int e = 0; e < ns; e++) {
        std::ostringstream el;
        const double *pr = g->shell_prop + GF_SHELL_NPROP * (size_t)e;
        int sec;
        auto it = thick.find(pr[3]);
        if (it != thick.end()) sec = it->second;
        else { sec = (int)thick.size() + 1; thick[pr[3]] = sec; }
        el << "Shell_1\t" << shell_num(e) << "\tMat\t" << mat_id(pr[0], pr[1], pr[2]) << "\tSec\t" << sec << "\tNodes";
        for (int a = 0; a < 6; a++) el << "\t" << g->shell_conn[6 * e + a] + 1;
        el << "\n";
        const int k = shell_num(e) - 1;
        if (k < 0 || k >= nb + ns || !lines[k].empty()) return fail("element numbers are not a permutation of 1..N");
        lines[k] = el.str();
    }
    std::ostringstream el;
    for (const std::string &ln : lines) el << ln;
    // node and element sets for loads and constraints
    std::ostringstream nsets, esets, lo, co;
    int n_nsets = 0, n_esets = 0;
    for (int l = 0; l < nl; l++) {
        int kind = g->load_kind[l];
        int a = g->load_set_off[l], b = g->load_set_off[l + 1];
        int r0 = g->load_tab_off[l], r1 = g->load_tab_off[l + 1];
        if (kind == GF_LOAD_SHELL_PRESSURE) {
            esets << "ElementSet\t" << ++n_esets << "\tElements\t" << b - a << "\tList";
            for (int q = a; q < b; q++) esets << "\t" << shell_num(g->load_set_items[q]);
            esets << "\n";
            lo << "ShellLoad\t" << l + 1 << "\tElementSet\t" << n_esets << "\tAreaUpdate\t"
               << (g->load_flags ? g->load_flags[l] : 0) << "\tNTimes\t" << r1 - r0 << "\n";
            for (int r = r0; r < r1; r++) lo << num(g->tab_time[r]) << "\t" << num(g->tab_val[6 * r]) << "\n";
        } else {
            nsets << "NodeSet\t" << ++n_nsets << "\tNodes\t" << b - a << "\tList";
            for (int q = a; q < b; q++) nsets << "\t" << g->load_set_items[q] + 1;
            nsets << "\n";
            lo << (kind == GF_LOAD_NODAL ? "NodalLoad" : "NodalFollowerLoad") << "\t" << l + 1 << "\tNodeSet\t"
               << n_nsets << "\tCS\t" << cs_id(g->load_cs + 9 * (size_t)l) << "\tNTimes\t" << r1 - r0 << "\n";
            for (int r = r0; r < r1; r++) {
                lo << num(g->tab_time[r]);
                for (int k = 0; k < 6; k++) lo << "\t" << num(g->tab_val[6 * r + k]);
                lo << "\n";
            }
        }
    }
    // constraints: one node set per distinct fixity pattern
    std::map<int, std::vector<int>> patterns;
    if (g->node_fix)
        for (int i = 0; i < nn; i++) {
            int pat = 0;
            for (int k = 0; k < 6; k++)
                if (g->node_fix[6 * i + k]) pat |= 1 << k;
            if (pat) patterns[pat].push_back(i + 1);
        }
    int nc = 0;
    static const char *dofs[6] = {"UX", "UY", "UZ", "ROTX", "ROTY", "ROTZ"};
    for (auto &pt : patterns) {
        nsets << "NodeSet\t" << ++n_nsets << "\tNodes\t" << pt.second.size() << "\tList";
        for (int id : pt.second) nsets << "\t" << id;
        nsets << "\n";
        co << "NodalConstraint\t" << ++nc << "\tNodeSet\t" << n_nsets << "\n";
        for (int k = 0; k < 6; k++) co << "\t" << dofs[k] << "\tBoolTable\t" << ((pt.first >> k) & 1) << "\n";
    }
    if (n_nsets) o << "NodeSets\t" << n_nsets << "\n" << nsets.str();
    if (n_esets) o << "ElementSets\t" << n_esets << "\n" << esets.str();
    o << "Elements\t" << nb + ns << "\n" << el.str();
    o << "Materials\t" << mats.size() << "\n";
    std::vector<std::pair<int, std::array<double, 3>>> mv;
    for (auto &m : mats) mv.push_back({m.second, m.first});
    std::sort(mv.begin(), mv.end());
    for (auto &m : mv)
        o << "Hooke\t" << m.first << "\tE\t" << num(m.second[0]) << "\tNu\t" << num(m.second[1]) << "\tRho\t"
          << num(m.second[2]) << "\n";
    if (!tubes.empty()) {
        o << "Sections\t" << tubes.size() << "\n";
        std::vector<std::pair<int, std::array<double, 3>>> tv;
        for (auto &t : tubes) tv.push_back({t.second, t.first});
        std::sort(tv.begin(), tv.end());
        for (auto &t : tv) {
            if (t.second[0] == 0.0)
                o << "Tube\t" << t.first << "\tDe\t" << num(t.second[1]) << "\tDi\t" << num(t.second[2]) << "\n";
            else
                o << "Rectangle\t" << t.first << "\tB\t" << num(t.second[1]) << "\tH\t" << num(t.second[2]) << "\n";
        }
    }
    if (!pipes.empty()) {
        o << "PipeSections\t" << pipes.size() << "\n";
        std::vector<std::pair<int, std::array<double, 6>>> pv;
        for (auto &t : pipes) pv.push_back({t.second, t.first});
        std::sort(pv.begin(), pv.end());
        for (auto &t : pv)
            o << "PS\t" << t.first << "\tEA\t" << num(t.second[0]) << "\tEI\t" << num(t.second[1]) << "\tGJ\t"
              << num(t.second[2]) << "\tGA\t" << num(t.second[3]) << "\tRho\t" << num(t.second[4])
              << "\tCDt\t0\tCDn\t0\tCAt\t0\tCAn\t0\tDe\t" << num(t.second[5]) << "\tDi\t0\n";
    }
    if (!thick.empty()) {
        o << "ShellSections\t" << thick.size() << "\n";
        std::vector<std::pair<int, double>> hv;
        for (auto &t : thick) hv.push_back({t.second, t.first});
        std::sort(hv.begin(), hv.end());
        for (auto &t : hv) o << "Homogeneous\t" << t.first << "\tThickness\t" << num(t.second) << "\n";
    }
    if (!css.empty()) {
        o << "CoordinateSystems\t" << css.size() << "\n";
        std::vector<std::pair<int, std::array<double, 6>>> cv;
        for (auto &c : css) cv.push_back({c.second, c.first});
        std::sort(cv.begin(), cv.end());
        for (auto &c : cv) {
            o << "CS\t" << c.first << "\tE1";
            for (int k = 0; k < 3; k++) o << "\t" << num(c.second[k]);
            o << "\tE3";
            for (int k = 3; k < 6; k++) o << "\t" << num(c.second[k]);
            o << "\n";
        }
    }
    if (nl) o << "Loads\t" << nl << "\n" << lo.str();
    if (nc) o << "Constraints\t" << nc << "\n" << co.str();
    const bool grav = g->gravity_on || g->gravity[0] != 0.0 || g->gravity[1] != 0.0 || g->gravity[2] != 0.0;
    if (grav || g->ocean_on) o << "Environment\n";
    if (grav)
        o << "GravityData\tG\t" << num(g->gravity[0]) << "\t" << num(g->gravity[1]) << "\t" << num(g->gravity[2])
          << "\tBoolTable\t" << (g->gravity_on ? 1 : 0) << "\n";
    if (g->ocean_on)
        o << "OceanData\nRhoFluid\t" << num(g->ocean_rho_fluid) << "\nSurfacePosition\t" << num(g->ocean_surface[0])
          << "\t" << num(g->ocean_surface[1]) << "\t" << num(g->ocean_surface[2]) << "\nSeaCurrent\tN\t0\tBoolTable\t0\n";
    o << "SolutionSteps\t1\nStatic\t1\nEndTime\t" << num(p->end_time) << "\nTimeStep\t" << num(p->time_step)
      << "\nMaxTimeStep\t" << num(p->max_time_step) << "\nMinTimeStep\t" << num(p->min_time_step) << "\nMaxIt\t"
      << p->max_it << "\nMinIt\t" << p->min_it << "\nConvIncrease\t" << p->conv_increase << "\nIncFactor\t"
      << num(p->inc_factor) << "\nSample\t" << p->sample << "\n";
    o << "ConvergenceCriteria\nForceTolerance\t" << num(p->force_tol) << "\nMomentTolerance\t" << num(p->moment_tol)
      << "\nForceMinimumReference\t" << num(p->force_min) << "\nMomentMinimumReference\t" << num(p->moment_min)
      << "\nConstraintMinimumReference\t" << num(p->constraint_min) << "\nDisplacementTolerance\t"
      << num(p->disp_tol) << "\nRotationTolerance\t" << num(p->rot_tol) << "\nLagrangeTolerance\t"
      << num(p->lag_tol) << "\nDisplacementMinimumReference\t" << num(p->disp_min)
      << "\nRotationMinimumReference\t" << num(p->rot_min) << "\nLagrangeMinimumReference\t" << num(p->lag_min)
      << "\nDivergenceReference\t" << num(p->divergence_ref) << "\n";
    o << "SolverOptions\nProcessors\t1\tLinSys\tDirect\n";
    std::ofstream f(path);
    if (!f) return fail(std::string("cannot write ") + path);
    f << o.str();
    return f ? GF_OK : fail(std::string("write failed: ") + path);
}
