// model.hpp - TetFIT input language: the .fdtd DEFINE/SET/END parser and the
// primitives it fills (material, bc, mesh, refinement, source, simulation,
// output, solid).  Behaviour follows the reference's parser
// (src/DGA/Discretization.hpp:45-211) and the SetParam methods of
// Material.hpp, BoundaryCondition.hpp, Mesh.hpp, Source.hpp, Simulation.hpp,
// Output.hpp and Solid.hpp.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace tfh {

// Constants of Utilities.hpp:92-95. PI is long double there; mu0 is computed
// from it in long double and rounded to double.
extern const long double PI_L;
extern const double MU0;
extern const double EPSILON0;
extern const double C0;

struct Error : std::runtime_error {
    explicit Error(const std::string& s) : std::runtime_error(s) {}
};

// Reference: MyThrow prints "Input file error at line N: <msg>" and throws.
[[noreturn]] void input_error(uint32_t line, const std::string& msg);

struct Vec3 {
    double x = 0, y = 0, z = 0;
    Vec3() = default;
    Vec3(double a, double b, double c) : x(a), y(b), z(c) {}
    double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline Vec3 operator+(const Vec3& a, const Vec3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(const Vec3& a, const Vec3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator*(const Vec3& a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec3 operator*(double s, const Vec3& a) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec3 operator/(const Vec3& a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline Vec3 operator-(const Vec3& a) { return {-a.x, -a.y, -a.z}; }
inline double dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(const Vec3& a, const Vec3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(const Vec3& a) { return std::sqrt(dot(a, a)); }

// Row-major 3x3 tensor.
struct Mat3 {
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static Mat3 diag(double a) { Mat3 r; r.m[0] = r.m[4] = r.m[8] = a; return r; }
    double& operator()(int i, int j) { return m[3 * i + j]; }
    double operator()(int i, int j) const { return m[3 * i + j]; }
    bool operator==(const Mat3& o) const {
        for (int i = 0; i < 9; ++i) if (m[i] != o.m[i]) return false;
        return true;
    }
    bool operator!=(const Mat3& o) const { return !(*this == o); }
    bool is_zero() const { for (double v : m) if (v != 0) return false; return true; }
    double det() const {
        return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
               m[2] * (m[3] * m[7] - m[4] * m[6]);
    }
    Mat3 inverse() const;
};
inline Vec3 operator*(const Mat3& A, const Vec3& v) {
    return {A.m[0] * v.x + A.m[1] * v.y + A.m[2] * v.z, A.m[3] * v.x + A.m[4] * v.y + A.m[5] * v.z,
            A.m[6] * v.x + A.m[7] * v.y + A.m[8] * v.z};
}

// Material.hpp:34-321. Tensors are eps0*eps_r etc.
struct Material {
    Mat3 epsilon = Mat3::diag(EPSILON0);
    Mat3 mu = Mat3::diag(MU0);
    Mat3 sigma, chi;
    bool pml = true;
    // Extension (no reference counterpart): `SET epsilon_random {lo,hi,seed}`
    // gives every Cartesian cell of this material its own isotropic eps_r,
    // uniform in [lo, hi), drawn from (seed, cell index) - per-cell random
    // media that the reference's input language cannot express.
    bool eps_random = false;
    double eps_lo = 1, eps_hi = 1;
    uint64_t eps_seed = 0;
    void set(uint32_t line, const std::string& p, const std::string& v);
};

// BoundaryCondition.hpp:34-136. `val` is 0 for pec/pmc; the reference leaves it
// uninitialised for a default-created bc (it is only read for pec/pmc edges).
struct BoundaryCondition {
    std::string type = "none";
    std::string dir = "none";
    double thickness = 0, val = 0;
    std::vector<uint32_t> surfs;
    void set(uint32_t line, const std::string& p, const std::string& v);
};

// Mesh.hpp:34-239.
struct MeshDef {
    std::string file = "none", name = "none", type = "none", mesher = "none";
    double scale = 1;
    double xmin = 0, xmax = 0, ymin = 0, ymax = 0, zmin = 0, zmax = 0;
    double xstep = 0, ystep = 0, zstep = 0;
    void set(uint32_t line, const std::string& p, const std::string& v, const std::string& base_dir);
};

struct Refinement {
    double x = 1, y = 1, z = 1;
    void set(uint32_t line, const std::string& p, const std::string& v);
};

// Source.hpp:34-583 (parameters; the evaluation lives in sources.cpp).
struct SourceDef {
    std::string st = " ", prof = "dc", carrier = "sin", dir = "x";
    double amp = 0, freq = 0, radius = 1, decay = 0, width = 1, delay = 0, impedance = 50;
    std::array<double, 3> kvec{{0, 0, 0}}, center{{0, 0, 0}}, location{{0, 0, 0}};
    std::array<std::string, 3> bfuncs{{"cos", "cos", "cos"}};
    std::vector<uint32_t> surface_label;
    void set(uint32_t line, const std::string& p, const std::string& v);
};

// Simulation.hpp:34-131. The source list always starts with {1} (:40).
struct SimulationDef {
    double duration = 0, courant = 1, step = 0, tol = 1e-8;
    std::vector<uint32_t> sources{1};
    uint32_t mesh = 1, output = 0;
    std::string method = "dga", solver = "none", analytic = "false";
    bool debug = false;
    void set(uint32_t line, const std::string& p, const std::string& v);
};

// Output.hpp:34-352. Probe coordinates, period and grid steps go through
// std::stof as in the reference (Output.hpp:108, :128, :154).
struct OutputDef {
    std::string mode = "none", name = "simulation";
    double period = 0, delay = 0;
    std::vector<Vec3> probes;
    double xstart = 0, ystart = 0, zstart = 0, xstop = 0, ystop = 0, zstop = 0;
    double xstep = 0, ystep = 0, zstep = 0;
    uint8_t ref_sys = 0;
    void set(uint32_t line, const std::string& p, const std::string& v);
};

// Solid.hpp:34-361 (box, sphere, cylinder inclusion rules).
struct SolidDef {
    std::string type = "none";
    uint32_t material = 0;
    double squareradius = 0;
    Vec3 center, corner, size;
    void set(uint32_t line, const std::string& p, const std::string& v);
    bool rule(double x, double y, double z) const;
};

// Parses "{a,b,c}" lists the way the reference's SetParam loops do: at most 3
// comma separated numbers, converted by `conv`.
std::vector<double> parse_brace_list(uint32_t line, const std::string& value, bool use_stof,
                                     bool allow_more = false);

struct Model {
    std::map<uint32_t, Material> materials;
    std::map<uint32_t, BoundaryCondition> bcs;
    std::map<uint32_t, MeshDef> meshes;
    std::map<uint32_t, Refinement> refinements;
    std::map<uint32_t, SourceDef> sources;
    std::map<uint32_t, SimulationDef> simulations;
    std::map<uint32_t, OutputDef> outputs;
    std::vector<uint32_t> bnd_to_bc = std::vector<uint32_t>(100, 0);  // BndToBC, D:61
    uint32_t method_line = 0;
    std::string base_dir;

    void load(const std::string& path);
};

// ReadSolidsFile, D:4219-4345.
std::map<uint32_t, SolidDef> read_solids_file(const std::string& path);

std::string resolve_path(const std::string& name, const std::string& base_dir);

}  // namespace tfh
