// model.cpp - .fdtd parser and primitive SetParam logic.
// Reference: src/DGA/Discretization.hpp:45-211 (parser), :4219-4345 (solids
// file), Material.hpp:51-308, BoundaryCondition.hpp:47-122, Mesh.hpp:47-239,
// Source.hpp:59-306, Simulation.hpp:52-97, Output.hpp:54-300, Solid.hpp:48-289.
#include "model.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unistd.h>

namespace tfh {

const long double PI_L = 3.141592653589793238L;
const double MU0 = (double)(4 * PI_L * 1e-7);
const double EPSILON0 = 8.854187817e-12;
const double C0 = 1 / std::sqrt(MU0 * EPSILON0);

void input_error(uint32_t line, const std::string& msg) {
    std::ostringstream os;
    os << "Input file error at line " << line << ": " << msg;
    throw Error(os.str());
}

Mat3 Mat3::inverse() const {
    // Adjugate / determinant (cofactor form of a 3x3 inverse).
    Mat3 r;
    double d = det();
    r.m[0] = (m[4] * m[8] - m[5] * m[7]) / d;
    r.m[1] = (m[2] * m[7] - m[1] * m[8]) / d;
    r.m[2] = (m[1] * m[5] - m[2] * m[4]) / d;
    r.m[3] = (m[5] * m[6] - m[3] * m[8]) / d;
    r.m[4] = (m[0] * m[8] - m[2] * m[6]) / d;
    r.m[5] = (m[2] * m[3] - m[0] * m[5]) / d;
    r.m[6] = (m[3] * m[7] - m[4] * m[6]) / d;
    r.m[7] = (m[1] * m[6] - m[0] * m[7]) / d;
    r.m[8] = (m[0] * m[4] - m[1] * m[3]) / d;
    return r;
}

static const std::vector<std::string> kDefinables = {"material", "source", "mesh", "bc",
                                                     "simulation", "geometry", "output", "refinement"};
static const std::vector<std::string> kSourceTypes = {"e", "b", "j", "h", "d"};
static const std::vector<std::string> kProfiles = {"wave", "gaussian", "const"};
static const std::vector<std::string> kCarriers = {"sin", "cos", "gaussian", "dc", "ricker", "gaussmod"};
static const std::vector<std::string> kDirections = {"x", "y", "z", "r", "psi", "rho", "phi", "theta"};
static const std::vector<std::string> kBcDirections = {"+x", "+y", "+z", "r", "+t", "rho", "-t", "-x", "-y", "-z"};
static const std::vector<std::string> kModes = {"sin", "cos"};
static const std::vector<std::string> kMeshTypes = {"tetrahedral", "cartesian", "none"};
static const std::vector<std::string> kMeshers = {"netgen", "gmsh", "none"};
static const std::vector<std::string> kOutputModes = {"silo", "probepoint", "maxerror", "l2norm"};
static const std::vector<std::string> kMethods = {"fdtd", "dga", "fem", "dgao2", "frac", "gfdtd", "fraco2", "fdtdo2"};
static const std::vector<std::string> kSolvers = {"cg", "agmg"};
static const std::vector<std::string> kSolidTypes = {"sphere", "box", "cylinder"};

static bool in(const std::vector<std::string>& v, const std::string& s) {
    return std::find(v.begin(), v.end(), s) != v.end();
}

std::vector<double> parse_brace_list(uint32_t line, const std::string& value, bool use_stof, bool allow_more) {
    std::vector<double> out;
    auto i = value.begin();
    if (i == value.end() || *i != '{') input_error(line, "coordinates must be inside braces {...,..,..}");
    ++i;
    while (i != value.end() && *i != ',' && *i != '}') {
        std::string coord;
        while (i != value.end() && *i != ',' && *i != '}') { coord.push_back(*i); ++i; }
        if (i == value.end()) input_error(line, "unbalanced bracket");
        if (out.size() >= 3 && !allow_more) input_error(line, "Maximum of three coordinates!");
        try {
            out.push_back(use_stof ? (double)std::stof(coord) : std::stod(coord));
        } catch (const std::exception&) {
            input_error(line, "invalid number '" + coord + "'");
        }
        if (*i == ',') ++i;
    }
    if (i == value.end()) input_error(line, "Undefined end to list of coordinates!");
    return out;
}

static double to_d(uint32_t line, const std::string& v) {
    try { return std::stod(v); } catch (const std::exception&) { input_error(line, "invalid number '" + v + "'"); }
}
static int to_i(uint32_t line, const std::string& v) {
    try { return std::stoi(v); } catch (const std::exception&) { input_error(line, "invalid integer '" + v + "'"); }
}

void Material::set(uint32_t line, const std::string& p, const std::string& v) {
    pml = false;
    auto diag_list = [&](Mat3& M, double scale) {
        auto l = parse_brace_list(line, v, false);
        for (size_t k = 0; k < l.size(); ++k) M(k, k) = scale * l[k];
    };
    if (p == "epsilon") epsilon = Mat3::diag(EPSILON0 * to_d(line, v));
    else if (p == "epsilon_xy") epsilon(0, 1) = epsilon(1, 0) = to_d(line, v);
    else if (p == "epsilon_xz") epsilon(0, 2) = epsilon(2, 0) = to_d(line, v);
    else if (p == "epsilon_yz") epsilon(1, 2) = epsilon(2, 1) = to_d(line, v);
    else if (p == "mu") mu = Mat3::diag(MU0 * to_d(line, v));
    else if (p == "mu_xy") mu(0, 1) = mu(1, 0) = to_d(line, v);
    else if (p == "mu_xz") mu(0, 2) = mu(2, 0) = to_d(line, v);
    else if (p == "mu_yz") mu(1, 2) = mu(2, 1) = to_d(line, v);
    else if (p == "sigma") sigma = Mat3::diag(to_d(line, v));
    else if (p == "sigma_xy") sigma(0, 1) = sigma(1, 0) = to_d(line, v);
    else if (p == "sigma_xz") sigma(0, 2) = sigma(2, 0) = to_d(line, v);
    else if (p == "sigma_yz") sigma(1, 2) = sigma(2, 1) = to_d(line, v);
    else if (p == "chi") chi = Mat3::diag(to_d(line, v));
    else if (p == "chi_xy") chi(0, 1) = chi(1, 0) = to_d(line, v);
    else if (p == "chi_xz") chi(0, 2) = chi(2, 0) = to_d(line, v);
    else if (p == "chi_yz") chi(1, 2) = chi(2, 1) = to_d(line, v);
    else if (p == "eps_vec") diag_list(epsilon, EPSILON0);
    else if (p == "mu_vec") diag_list(mu, MU0);
    else if (p == "sigma_vec") diag_list(sigma, 1.0);
    else if (p == "chi_vec") diag_list(chi, 1.0);
    else if (p == "epsilon_random") {
        auto l = parse_brace_list(line, v, false);
        if (l.size() != 3 || !(l[0] > 0) || !(l[1] >= l[0]) || l[2] < 0)
            input_error(line, "epsilon_random expects {lo,hi,seed} with 0 < lo <= hi, seed >= 0");
        eps_random = true;
        eps_lo = l[0];
        eps_hi = l[1];
        eps_seed = (uint64_t)l[2];
        epsilon = Mat3::diag(EPSILON0 * eps_lo);
    }
    else input_error(line, "Unrecognized material parameter! Available: epsilon, mu, sigma, chi");
}

void BoundaryCondition::set(uint32_t line, const std::string& p, const std::string& v) {
    if (p == "type") {
        if (v == "pml") type = v;
        else if (v == "pmc" || v == "pec") { type = v; val = 0; }
        else input_error(line, "Unrecognized boundary condition type! Available: pec, pmc, pml");
    } else if (p == "direction") {
        if (!in(kBcDirections, v)) input_error(line, "Unrecognized direction! Available: +-x, +-y, +-z, r, +-t, rho");
        dir = v;
    } else if (p == "width") thickness = to_d(line, v);
    else if (p == "surface") surfs.push_back((uint32_t)to_i(line, v));
    else if (p == "surfaces") {
        // BoundaryCondition.hpp:73-109 accepts any number of labels.
        auto l = parse_brace_list(line, v, false, true);
        for (double x : l) surfs.push_back((uint32_t)x);
    } else input_error(line, "Unrecognized boundary condition parameter! Available: type");
}

std::string resolve_path(const std::string& name, const std::string& base_dir) {
    if (access(name.c_str(), R_OK) == 0) return name;
    if (!name.empty() && name[0] != '/' && !base_dir.empty()) {
        std::string alt = base_dir + "/" + name;
        if (access(alt.c_str(), R_OK) == 0) return alt;
    }
    return "";
}

void MeshDef::set(uint32_t line, const std::string& p, const std::string& v, const std::string& base_dir) {
    auto grid = [&](double& a, double& s, double& b) {
        auto l = parse_brace_list(line, v, false);
        if (l.size() > 0) a = l[0];
        if (l.size() > 1) s = l[1];
        if (l.size() > 2) b = l[2];
        if (!(a <= b) || !(s > 0)) input_error(line, "grid needs start <= stop and step > 0");
    };
    if (p == "file") {
        std::string r = resolve_path(v, base_dir);
        if (r.empty()) input_error(line, "Mesh input file does not exist or it is not in specified path");
        file = r;
    } else if (p == "name") name = v;
    else if (p == "type") {
        if (!in(kMeshTypes, v)) input_error(line, "Undefined mesh type! Available: tetrahedral, cartesian");
        type = v;
    } else if (p == "mesher") {
        if (!in(kMeshers, v)) input_error(line, "Undefined mesher! Available: netgen, gmsh");
        mesher = v;
    } else if (p == "xgrid") grid(xmin, xstep, xmax);
    else if (p == "ygrid") grid(ymin, ystep, ymax);
    else if (p == "zgrid") grid(zmin, zstep, zmax);
    else if (p == "scalefactor") scale = to_d(line, v);
    else input_error(line, "Undefined mesh parameter! Available: file, name, type, mesher, scalefactor");
}

void Refinement::set(uint32_t line, const std::string& p, const std::string& v) {
    if (p == "vector") {
        auto l = parse_brace_list(line, v, false);
        if (l.size() > 0) x = l[0];
        if (l.size() > 1) y = l[1];
        if (l.size() > 2) z = l[2];
    } else if (p == "x") x = to_d(line, v);
    else if (p == "y") y = to_d(line, v);
    else if (p == "z") z = to_d(line, v);
    else input_error(line, "Undefined refinement parameter! Available: x, y, z, vector");
    if (x < 1 || y < 1 || z < 1) input_error(line, "refinement factors must be >= 1");
}

void SourceDef::set(uint32_t line, const std::string& p, const std::string& v) {
    if (p == "direction") {
        if (!in(kDirections, v)) input_error(line, "Unrecognized direction! Available: x, y, z");
        dir = v;
    } else if (p == "type") {
        if (!in(kSourceTypes, v)) input_error(line, "Unrecognized source type!Available: e (electric field), b (magnetic field), j (current density)");
        st = v;
    } else if (p == "profile") {
        if (!in(kProfiles, v)) input_error(line, "Unrecognized source profile! Available: const, wave, gaussian");
        prof = v;
    } else if (p == "carrier") {
        if (!in(kCarriers, v)) input_error(line, "Unrecognized source carrier! Available: dc, sin, cos, gaussian, ricker, gaussmod");
        carrier = v;
    } else if (p == "center" || p == "wavevector" || p == "location") {
        auto l = parse_brace_list(line, v, false);
        auto& dst = p == "center" ? center : (p == "wavevector" ? kvec : location);
        for (size_t k = 0; k < l.size(); ++k) dst[k] = l[k];
    } else if (p == "mode") {
        // Source.hpp:199-237: spaces are dropped inside the braces.
        auto i = v.begin();
        if (i == v.end() || *i != '{') input_error(line, "coordinates must be inside braces {...,..,..}");
        ++i;
        int k = 0;
        while (i != v.end() && *i != ',' && *i != '}') {
            std::string c;
            while (i != v.end() && *i != ',' && *i != '}') { if (*i != ' ') c.push_back(*i); ++i; }
            if (i == v.end()) input_error(line, "unbalanced bracket");
            if (k >= 3) input_error(line, "Maximum of three coordinates!");
            bfuncs[k++] = c;
            if (*i == ',') ++i;
        }
        if (i == v.end()) input_error(line, "Undefined end to list of coordinates!");
    } else if (p == "Xmode" || p == "Ymode" || p == "Zmode") {
        if (!in(kModes, v)) input_error(line, "Unrecognized base function! Available: sin, cos");
        bfuncs[p[0] - 'X'] = v;
    } else if (p == "amplitude") amp = to_d(line, v);
    else if (p == "delay") delay = to_d(line, v);
    else if (p == "impedance") impedance = to_d(line, v);
    else if (p == "frequency") freq = to_d(line, v);
    else if (p == "radius") radius = to_d(line, v);
    else if (p == "decay") decay = to_d(line, v);
    else if (p == "width") width = to_d(line, v);
    else if (p == "kx") kvec[0] = to_d(line, v);
    else if (p == "ky") kvec[1] = to_d(line, v);
    else if (p == "kz") kvec[2] = to_d(line, v);
    else if (p == "surface") surface_label.push_back((uint32_t)to_i(line, v));
    else input_error(line, "unrecognized parameter for source!");
}

void SimulationDef::set(uint32_t line, const std::string& p, const std::string& v) {
    if (p == "source") {
        sources.push_back((uint32_t)to_d(line, v));
        std::sort(sources.begin(), sources.end());
        sources.erase(std::unique(sources.begin(), sources.end()), sources.end());
    } else if (p == "mesh") mesh = (uint32_t)to_d(line, v);
    else if (p == "duration") duration = to_d(line, v);
    else if (p == "courant") courant = to_d(line, v);
    else if (p == "step") step = to_d(line, v);
    else if (p == "output") output = (uint32_t)to_d(line, v);
    else if (p == "method") {
        if (!in(kMethods, v)) input_error(line, "Undefined simulation method! Available: fdtd, dga, fem");
        method = v;
    } else if (p == "solver") {
        if (!in(kSolvers, v)) input_error(line, "Unavailable solver! Available: cg, agmg");
        solver = v;
    } else if (p == "tolerance") tol = to_d(line, v);
    else if (p == "analytic") analytic = v;
    else if (p == "debug") {
        if (v == "true") debug = true;
        else if (v == "false") debug = false;
        else input_error(line, "Undefined simulation parameter! Available: source, mesh, duration, output");
    } else input_error(line, "Undefined simulation parameter! Available: source, mesh, duration, output");
}

void OutputDef::set(uint32_t line, const std::string& p, const std::string& v) {
    auto grid = [&](double& a, double& s, double& b) {
        auto l = parse_brace_list(line, v, true);
        if (l.size() > 0) a = l[0];
        if (l.size() > 1) s = l[1];
        if (l.size() > 2) b = l[2];
        if (!(a <= b) || !(s > 0)) input_error(line, "grid needs start <= stop and step > 0");
    };
    if (p == "mode") {
        if (!in(kOutputModes, v)) input_error(line, "Undefined output mode type! Available: silo, probepoint, maxerror, l2norm");
        mode = v;
    } else if (p == "radiator") {
        // radiator surfaces are not used by the time loops
    } else if (p == "delay") delay = to_d(line, v);
    else if (p == "coordinates") {
        if (v == "cartesian") ref_sys = 0;
        else if (v == "cylindrical") ref_sys = 1;
        else if (v == "spherical") ref_sys = 2;
        else input_error(line, "Invalid system of coordinates! Available: cartesian, cylindrical, spherical");
    } else if (p == "probe") {
        auto l = parse_brace_list(line, v, true);
        Vec3 pv;
        for (size_t k = 0; k < l.size(); ++k) pv[(int)k] = l[k];
        probes.push_back(pv);
    } else if (p == "period") {
        try { period = std::stof(v); } catch (const std::exception&) { input_error(line, "invalid number"); }
    } else if (p == "xgrid") grid(xstart, xstep, xstop);
    else if (p == "ygrid") grid(ystart, ystep, ystop);
    else if (p == "zgrid") grid(zstart, zstep, zstop);
    else if (p == "grid") {
        if (v == "on") {
            // Output.hpp:286-296: float-rounded steps accumulated in double.
            for (double z = zstart; z <= zstop; z += zstep)
                for (double y = ystart; y <= ystop; y += ystep)
                    for (double x = xstart; x <= xstop; x += xstep) probes.push_back(Vec3(x, y, z));
        } else if (v != "off") input_error(line, "Undefined output parameter! Available: mode, period, probe, name");
    } else if (p == "error") {
        if (v != "on" && v != "off") input_error(line, "Undefined output parameter! Available: mode, period, probe, name");
    } else if (p == "name") name = v;
    else input_error(line, "Undefined output parameter! Available: mode, period, probe, name");
}

void SolidDef::set(uint32_t line, const std::string& p, const std::string& v) {
    if (p == "type") {
        if (!in(kSolidTypes, v)) input_error(line, "Unavailable solid type! Available: sphere, box, cylinder");
        type = v;
    } else if (p == "material") material = (uint32_t)std::atoi(v.c_str());
    else if (p == "radius") {
        double r = to_d(line, v);
        if (r <= 0) input_error(line, "Positive definite quantity forced to negative value!");
        squareradius = r * r;
    } else if (p == "center" || p == "corner" || p == "size") {
        auto l = parse_brace_list(line, v, false);
        Vec3& dst = p == "center" ? center : (p == "corner" ? corner : size);
        for (size_t k = 0; k < l.size(); ++k) {
            if (p == "size" && l[k] < 0) input_error(line, "Positive definite quantity forced to negative value!");
            dst[(int)k] = l[k];
        }
    } else input_error(line, "Unrecognized solid parameter! Available: type, radius, center, corner, size");
}

bool SolidDef::rule(double x, double y, double z) const {
    Vec3 p(x, y, z);
    if (type == "sphere") {
        Vec3 d = p - center;
        return dot(d, d) <= squareradius;
    } else if (type == "box") {
        Vec3 d = p - corner;
        for (int k = 0; k < 3; ++k)
            if (d[k] > size[k] || d[k] < 0) return false;
        return true;
    } else if (type == "cylinder") {
        double ss = dot(size, size);
        if (std::sqrt(ss) == 0) return false;
        Vec3 pd = p - center;
        double dt = dot(pd, size);
        if (dt < -1e-12 || dt > ss + 1e-12) return false;
        double dsq = dot(pd, pd) - dt * dt / ss;
        return dsq <= squareradius + 1e-12;
    }
    return false;
}

// One line of the DEFINE/SET/END language, split like sscanf("%s %s %s").
// Tokens missing from a short line keep the previous line's values, as the
// reference's fixed char buffers do.
struct LineTokens {
    std::string action, token, value;
    void scan(const std::string& line) {
        std::istringstream is(line);
        std::string a, t, v;
        if (is >> a) action = a; else return;
        if (is >> t) token = t; else return;
        if (is >> v) value = v;
    }
};

// Extracts "{...}" from the raw line when the value starts with '{' (D:144-161).
static std::string brace_value(uint32_t line_no, const std::string& line) {
    size_t a = line.find('{');
    size_t b = line.find('}', a == std::string::npos ? 0 : a);
    if (a == std::string::npos) input_error(line_no, "unbalanced bracket");
    if (b == std::string::npos) input_error(line_no, "unbalanced bracket");
    return line.substr(a, b - a + 1);
}

void Model::load(const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) throw Error("Input file missing! Correct use is: \"tetfit input_file\" (" + path + ")");
    size_t slash = path.rfind('/');
    base_dir = slash == std::string::npos ? std::string(".") : path.substr(0, slash);

    std::string line, thing;
    uint32_t label = 0, line_no = 1;
    bool in_def = false;
    LineTokens tk;
    while (std::getline(f, line)) {
        if (!line.empty() && line[0] != '#' && line[0] != '\n' && line[0] != '\r') {
            tk.scan(line);
            const std::string& instr = tk.action;
            if (!instr.empty() && instr[0] != '#') {
                if (!in_def) {
                    if (instr != "DEFINE") input_error(line_no, "define something before setting variables");
                    if (!in(kDefinables, tk.token))
                        input_error(line_no, "can only define one of the following: material, mesh, refinement, boundary condition, simulation, output, source");
                    thing = tk.token;
                    label = (uint32_t)to_i(line_no, tk.value);
                    in_def = true;
                    if (thing == "material") materials[label] = Material();
                    else if (thing == "bc") bcs[label] = BoundaryCondition();
                    else if (thing == "mesh") meshes[label] = MeshDef();
                    else if (thing == "refinement") refinements[label] = Refinement();
                    else if (thing == "source") sources[label] = SourceDef();
                    else if (thing == "simulation") simulations[label] = SimulationDef();
                    else if (thing == "output") outputs[label] = OutputDef();
                } else if (instr == "END") {
                    if (tk.token != thing || std::atoi(tk.value.c_str()) != (int)label)
                        input_error(line_no, "ending unopened definition");
                    in_def = false;
                    if (thing == "bc") {
                        for (uint32_t s : bcs[label].surfs) {
                            if (s >= bnd_to_bc.size()) input_error(line_no, "surface labels must be < 100");
                            bnd_to_bc[s] = label;
                        }
                    }
                } else if (instr != "SET") {
                    input_error(line_no, "Unknown instruction inside define block");
                } else {
                    std::string val = tk.value;
                    if (!tk.value.empty() && tk.value[0] == '{') val = brace_value(line_no, line);
                    if (thing == "material") materials[label].set(line_no, tk.token, val);
                    else if (thing == "bc") bcs[label].set(line_no, tk.token, val);
                    else if (thing == "mesh") meshes[label].set(line_no, tk.token, val, base_dir);
                    else if (thing == "refinement") refinements[label].set(line_no, tk.token, val);
                    else if (thing == "source") sources[label].set(line_no, tk.token, val);
                    else if (thing == "simulation") {
                        simulations[label].set(line_no, tk.token, val);
                        if (tk.token == "method") method_line = line_no;
                    } else if (thing == "output") outputs[label].set(line_no, tk.token, val);
                    else input_error(line_no, "can only define one of the following: material, mesh, refinement, boundary condition, simulation, output, source");
                }
            }
        }
        ++line_no;
    }
    if (in_def) input_error(line_no, "File ended unexpectedly before ending definition");
}

std::map<uint32_t, SolidDef> read_solids_file(const std::string& path) {
    std::map<uint32_t, SolidDef> solids;
    std::ifstream f(path);
    if (path.empty() || !f.is_open()) throw Error("Invalid mesh file name");
    const std::string pre = "In cartesian mesh definition file: ";
    std::string line, thing;
    uint32_t label = 0, line_no = 1;
    bool in_def = false;
    LineTokens tk;
    try {
        while (std::getline(f, line)) {
            if (!line.empty() && line[0] != '#' && line[0] != '\n' && line[0] != '\r') {
                tk.scan(line);
                if (!tk.action.empty() && tk.action[0] != '#') {
                    if (!in_def) {
                        if (tk.action != "DEFINE") input_error(line_no, "define something before setting variables");
                        if (tk.token != "grid" && tk.token != "solid")
                            input_error(line_no, "Unknown primivite: can only define solid inside cartesian meshes");
                        thing = tk.token;
                        label = (uint32_t)to_d(line_no, tk.value);
                        in_def = true;
                        if (thing == "solid") solids[label] = SolidDef();
                    } else if (tk.action == "END") {
                        if (tk.token != thing || std::atoi(tk.value.c_str()) != (int)label)
                            input_error(line_no, "ending unopened definition");
                        in_def = false;
                    } else if (tk.action != "SET") {
                        input_error(line_no, "Unknown instruction inside define block");
                    } else {
                        std::string val = tk.value;
                        if (!tk.value.empty() && tk.value[0] == '{') val = brace_value(line_no, line);
                        if (thing == "solid") solids[label].set(line_no, tk.token, val);
                        else input_error(line_no, "Unknown primivite: can only define solid inside cartesian meshes");
                    }
                }
            }
            ++line_no;
        }
        if (in_def) input_error(line_no, "File ended unexpectedly before ending definition");
    } catch (const Error& e) {
        throw Error(pre + e.what());
    }
    return solids;
}

}  // namespace tfh
