// fit.cpp - cartesian mesher and FIT operator producer (method fdtd).
// Restates ReadStructuredMesh (src/DGA/Discretization.hpp:4347-5690):
// cube-by-cube construction in k, j, i order with reuse of the bottom (z-),
// left (y-) and back (x-) neighbours' entities, signed incidences C_vec /
// Ct_vec / curl / dual_curl, material averages, edge classes
// (common / uncommon / bc / src) and the FIT coefficients M_h, M_q, M_mu,
// M_nu. Also the FIT probe location (FindFitProbe D:3590-3643), probe weights
// (GetFitElectricField D:3555, GetFitMagneticField D:3570) and the boundary
// dual edge vectors of h sources (ComputeHfieldSource D:4148-4199).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "build.hpp"

namespace tfh {

namespace {

// Local numbering of a cube (D:4643-4934):
//   nodes 0..7 at offset (p&1)x + (p>>1&1)y + (p>>2&1)z
//   faces 0 z-, 1 y-, 2 x-, 3 x+, 4 y+, 5 z+
//   edges 0 (P0P1,x) 1 (P0P2,y) 2 (P0P4,z) 3 (P1P3,y) 4 (P1P5,z) 5 (P2P3,x)
//         6 (P2P6,z) 7 (P3P7,z) 8 (P4P5,x) 9 (P4P6,y) 10 (P5P7,y) 11 (P6P7,x)
const int kEdgeNodes[12][2] = {{0, 1}, {0, 2}, {0, 4}, {1, 3}, {1, 5}, {2, 3},
                               {2, 6}, {3, 7}, {4, 5}, {4, 6}, {5, 7}, {6, 7}};
const int kEdgeAxis[12] = {0, 1, 2, 1, 2, 0, 2, 2, 0, 1, 1, 0};
// dual_curl pushed for a new edge: +1 for x and z edges, -1 for y (D:4960-5080)
const int kDualCurl[3] = {1, -1, 1};
const int kFaceEdges[6][4] = {{0, 1, 3, 5}, {0, 2, 4, 8}, {1, 2, 6, 9},
                              {3, 4, 7, 10}, {5, 6, 7, 11}, {8, 9, 10, 11}};
const int kFaceSigns[6][4] = {{1, -1, 1, -1}, {-1, 1, -1, 1}, {1, -1, 1, -1},
                              {1, -1, 1, -1}, {-1, 1, -1, 1}, {1, -1, 1, -1}};
const int kFaceCurl[6] = {1, -1, 1, 1, -1, 1};
const int kFaceAxis[6] = {2, 1, 0, 0, 1, 2};      // normal axis
const int kFaceLabel[6] = {1, 5, 2, 4, 3, 6};     // boundary label at creation
// Reuse from neighbours, in priority order bottom > left > back:
// (neighbour kind 0 bottom / 1 left / 2 back, position in the neighbour).
struct Src { int nb, pos; };
const std::vector<Src> kEdgeReuse[12] = {
    {{0, 8}, {1, 5}}, {{0, 9}, {2, 3}}, {{1, 6}, {2, 4}}, {{0, 10}}, {{1, 7}}, {{0, 11}},
    {{2, 7}},         {},               {{1, 11}},        {{2, 10}}, {},        {}};
const std::vector<Src> kNodeReuse[8] = {{{0, 4}, {1, 2}, {2, 1}}, {{0, 5}, {1, 3}}, {{0, 6}, {2, 3}}, {{0, 7}},
                                        {{1, 6}, {2, 5}},         {{1, 7}},         {{2, 7}},         {}};
const std::vector<Src> kFaceReuse[6] = {{{0, 5}}, {{1, 4}}, {{2, 3}}, {}, {}, {}};

// A table of n values filled on all host threads (the mesher's per-entity
// tables take several GB at 256^3 cells; a sequential zero fill of them cost
// more than building them).
template <class T>
struct ParTable {
    std::unique_ptr<T[]> p;
    size_t n = 0;
    ParTable() = default;
    ParTable(size_t n_, T v) : p(new T[n_]), n(n_) {  // trivial T: left uninitialised by new
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n_; ++i) p[i] = v;
    }
    T& operator[](size_t i) { return p[i]; }
    const T& operator[](size_t i) const { return p[i]; }
    size_t size() const { return n; }
    T* begin() { return p.get(); }
    T* end() { return p.get() + n; }
};

// The 12 E-probe weights use the dual area of each edge's axis (D:3555-3568).
}  // namespace

void build_fit(const Model& m, const SimulationDef& sim, Build& b) {
    PhaseClock clk;
    auto mit = m.meshes.find(sim.mesh);
    if (mit == m.meshes.end()) throw Error("simulation refers to an undefined mesh");
    const MeshDef& msh = mit->second;
    std::map<uint32_t, SolidDef> solids = read_solids_file(msh.file);
    // ---- PML (D:4414-4556): a bc of type pml adds floor(width scale / L)
    // cells on its side of the grid; the inner box [xmin, xmax] keeps its
    // place, and every solid gets an extension solid (new labels after the
    // last, taken in reverse label order) grown by the widths
    // (Solid::ConstructPMLExtension, Solid.hpp:297-354), written to text at
    // the stream's default 6 significant digits and read back. Only the +x,
    // +y, +z sides are supported: the reference's mesher links a cube to its
    // lower neighbours only above the inner box's lower corner (D:4621-4633),
    // so layers on the -x, -y, -z sides (and the rho / r / t forms, which
    // grow every side, or index past a 2-entry width table) come out
    // disconnected; cylinders with PML trip an assertion on that table.
    double pml_width[3] = {0, 0, 0};  // +x, +y, +z
    bool pml_present = false;
    for (auto& kv : m.bcs) {
        const BoundaryCondition& bc = kv.second;
        if (bc.type != "pml") continue;
        pml_present = true;
        if (!bc.surfs.empty()) throw Error("a pml boundary condition takes a direction and a width, not surfaces");
        if (bc.dir == "+x") pml_width[0] = std::fabs(bc.thickness);
        else if (bc.dir == "+y") pml_width[1] = std::fabs(bc.thickness);
        else if (bc.dir == "+z") pml_width[2] = std::fabs(bc.thickness);
        else
            throw Error("pml direction '" + bc.dir + "' is not supported (only +x, +y, +z: the reference's mesh "
                        "is disconnected in layers on the lower sides, D:4621-4633)");
    }
    if (pml_present) {
        auto g6 = [](double v) {
            char buf[64];
            std::snprintf(buf, sizeof buf, "%g", v);
            return std::stod(buf);
        };
        uint32_t label = solids.empty() ? 0 : solids.rbegin()->first;
        std::vector<std::pair<uint32_t, SolidDef>> ext;
        for (auto it = solids.rbegin(); it != solids.rend(); ++it) {
            SolidDef s = it->second;
            if (s.type == "cylinder")
                throw Error("cylinder solids with PML are not supported (the reference asserts on a 2-entry "
                            "width table, D:4416 with Solid.hpp:335)");
            if (s.type == "box") {
                for (int a = 0; a < 3; ++a) {
                    s.corner[a] = g6(s.corner[a]);
                    s.size[a] = g6(s.size[a] + pml_width[a]);
                }
            } else if (s.type == "sphere") {
                for (int a = 0; a < 3; ++a) s.center[a] = g6(s.center[a]);
                const double r = g6(std::sqrt(s.squareradius));
                s.squareradius = r * r;
            }
            ext.push_back({++label, s});
        }
        for (auto& e : ext) solids[e.first] = e.second;
    }
    std::map<uint32_t, Material> materials = m.materials;  // operator[] may insert defaults
    std::map<uint32_t, BoundaryCondition> bcs = m.bcs;
    std::map<uint32_t, SourceDef> sources = m.sources;

    // Material adjacency labels for interface faces (D:330-347).
    uint32_t maxmat = materials.empty() ? 0 : materials.rbegin()->first;
    std::vector<uint32_t> matad((size_t)(maxmat + 1) * (maxmat + 1), 0);
    {
        uint32_t lab = 6;
        for (uint32_t i = 0; i <= maxmat; ++i)
            for (uint32_t j = i + 1; j <= maxmat; ++j) {
                matad[i * (maxmat + 1) + j] = ++lab;
                matad[j * (maxmat + 1) + i] = lab;
            }
    }
    auto matad_at = [&](uint32_t a, uint32_t c) -> uint32_t {
        if (a > maxmat || c > maxmat) return 0;
        return matad[a * (maxmat + 1) + c];
    };

    const double scale = msh.scale;
    const double xmin = scale * msh.xmin, ymin = scale * msh.ymin, zmin = scale * msh.zmin;
    const double xmax = scale * msh.xmax, ymax = scale * msh.ymax, zmax = scale * msh.zmax;
    double Lx = scale * msh.xstep, Ly = scale * msh.ystep, Lz = scale * msh.zstep;
    for (auto& r : m.refinements) { Lx /= r.second.x; Ly /= r.second.y; Lz /= r.second.z; }
    const uint32_t Npml[3] = {(uint32_t)(pml_width[0] * scale / Lx), (uint32_t)(pml_width[1] * scale / Ly),
                              (uint32_t)(pml_width[2] * scale / Lz)};
    const uint32_t Nx = (uint32_t)(std::fabs(xmax - xmin) / Lx) + Npml[0];
    const uint32_t Ny = (uint32_t)(std::fabs(ymax - ymin) / Ly) + Npml[1];
    const uint32_t Nz = (uint32_t)(std::fabs(zmax - zmin) / Lz) + Npml[2];
    const double true_max[3] = {xmax + double(Npml[0]) * Lx, ymax + double(Npml[1]) * Ly, zmax + double(Npml[2]) * Lz};
    if (Nx == 0 || Ny == 0 || Nz == 0) throw Error("cartesian grid has no cells");
    const uint64_t totE = (uint64_t)Nx * (Ny + 1) * (Nz + 1) + (uint64_t)(Nx + 1) * Ny * (Nz + 1) +
                          (uint64_t)(Nx + 1) * (Ny + 1) * Nz;
    const uint64_t totF = (uint64_t)Nx * Ny * (Nz + 1) + (uint64_t)Nz * Nx * (Ny + 1) + (uint64_t)Ny * Nz * (Nx + 1);
    const uint64_t totV = (uint64_t)Nx * Ny * Nz;
    const uint64_t totP = (uint64_t)(Nx + 1) * (Ny + 1) * (Nz + 1);
    if (totE >= (1ull << 31) || totF >= (1ull << 31)) throw Error("cartesian grid too large for int32 indices");

    b.L[0] = Lx; b.L[1] = Ly; b.L[2] = Lz;
    const double area[3] = {Ly * Lz, Lx * Lz, Lx * Ly};  // area_x, area_y, area_z
    const double da[3] = {area[0] / 4, area[1] / 4, area[2] / 4};
    const double len[3] = {Lx, Ly, Lz};

    std::vector<int32_t>& E_cluster = b.e_cluster;  // 12 per cube
    std::vector<int32_t>& D = b.d_faces;            // 6 per cube
    ParTable<int32_t> P_cluster(totV * 8, 0);       // 8 per cube
    E_cluster.assign(totV * 12, 0);
    D.assign(totV * 6, 0);
    b.pts.assign(totP * 3, 0.0);
    ParTable<int32_t> Dt(totF * 2, -1);   // signed cube ids, up to 2 per face: value = cube, sign separate
    ParTable<int8_t> Dt_sgn(totF * 2, 0);
    ParTable<uint8_t> Dt_n(totF, 0);
    std::vector<int32_t>& G = b.g_nodes;  // 2 per edge
    G.assign(totE * 2, -1);
    // grid positions of edges and faces (tf_fit_ops edge_sid / face_sid)
    b.cells[0] = (int32_t)Nx; b.cells[1] = (int32_t)Ny; b.cells[2] = (int32_t)Nz;
    b.edge_sid.assign(totE, -1);
    b.face_sid.assign(totF, -1);
    const int64_t nEx = (int64_t)Nx * (Ny + 1) * (Nz + 1), nEy = (int64_t)(Nx + 1) * Ny * (Nz + 1);
    const int64_t nFx = (int64_t)(Nx + 1) * Ny * Nz, nFy = (int64_t)Nx * (Ny + 1) * Nz;
    auto edge_pos = [&](int a, int64_t i, int64_t j, int64_t k) -> int32_t {
        if (a == 0) return (int32_t)((k * (Ny + 1) + j) * Nx + i);
        if (a == 1) return (int32_t)(nEx + (k * Ny + j) * (Nx + 1) + i);
        return (int32_t)(nEx + nEy + (k * (Ny + 1) + j) * (Nx + 1) + i);
    };
    auto face_pos = [&](int a, int64_t i, int64_t j, int64_t k) -> int32_t {
        if (a == 0) return (int32_t)((k * Ny + j) * (Nx + 1) + i);
        if (a == 1) return (int32_t)(nFx + (k * (Ny + 1) + j) * Nx + i);
        return (int32_t)(nFx + nFy + (k * Ny + j) * Nx + i);
    };
    b.c_edge.assign(totF * 4, 0);
    ParTable<int32_t> ct_face(totE * 4, 0);
    ParTable<int8_t> ct_sgn(totE * 4, 0);
    ParTable<uint8_t> ct_n(totE, 0);
    b.curl.assign(totF, 1);
    b.dual_curl.assign(totE, 0);
    ParTable<double> edge_len(totE, 0), face_area(totF, 0);
    ParTable<double> avg_eps(totE, 0), avg_sigma(totE, 0), avg_ni(totF, 0), avg_mag_sigma(totF, 0);
    ParTable<uint8_t> ele_lossy(totE, 0), mag_lossy(totF, 0);
    ParTable<uint32_t> boundary_face(totF, 0);
    ParTable<uint8_t> tbc(totF, 0);
    // Material of every cube (WhichSolid D:10110: the first solid by label
    // that holds the cube's centre), on all host threads. The corners are the
    // running sums of the steps (the reference's cube loop adds the step
    // once per cube), so the solid tests see the same coordinates bit for bit.
    std::vector<double> xs(Nx), ys(Ny), zs(Nz);
    {
        double v = xmin;
        for (uint32_t i = 0; i < Nx; ++i) { xs[i] = v; v += Lx; }
        v = ymin;
        for (uint32_t j = 0; j < Ny; ++j) { ys[j] = v; v += Ly; }
        v = zmin;
        for (uint32_t k = 0; k < Nz; ++k) { zs[k] = v; v += Lz; }
    }
    std::vector<const SolidDef*> solid_list;
    for (auto& s : solids) solid_list.push_back(&s.second);
    if (solid_list.size() >= 65535) throw Error("cartesian mesher: too many solids");
    ParTable<uint32_t> vol_material(totV, 0);
    ParTable<uint16_t> vol_solid(totV, 0);  // 0: no solid, s + 1: solid_list[s]
    std::vector<uint8_t> solid_used(solid_list.size() + 1, 0);
    const int64_t nrow = (int64_t)Ny * Nz;  // grid rows (j, k), row = k Ny + j
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < nrow; ++row) {
        const uint32_t j = (uint32_t)(row % Ny), k = (uint32_t)(row / Ny);
        for (uint32_t i = 0; i < Nx; ++i) {
            uint32_t mat = 0, sid = 0;
            for (size_t q = 0; q < solid_list.size(); ++q)
                if (solid_list[q]->rule(xs[i] + 0.5 * Lx, ys[j] + 0.5 * Ly, zs[k] + 0.5 * Lz)) {
                    mat = solid_list[q]->material;
                    sid = (uint32_t)q + 1;
                    break;
                }
            vol_material[(size_t)row * Nx + i] = mat;
            vol_solid[(size_t)row * Nx + i] = (uint16_t)sid;
            __atomic_store_n(&solid_used[sid], (uint8_t)1, __ATOMIC_RELAXED);
        }
    }
    // the material of each solid that holds a cube (operator[] inserts the
    // default material for an undefined label, as the reference's lookup does)
    std::vector<const Material*> solid_mat(solid_list.size() + 1, nullptr);
    for (size_t q = 0; q <= solid_list.size(); ++q)
        if (solid_used[q]) solid_mat[q] = &materials[q ? solid_list[q - 1]->material : 0u];

    // ---- numbering (D:4621-4934) on all host threads. The reference walks
    // the cubes in k, j, i order, takes a face / edge / node from the bottom,
    // left or back neighbour when that neighbour exists and numbers it next
    // otherwise. Which local entities a cube numbers thus depends only on
    // which of its three neighbours exist: each grid row's first numbers are a
    // prefix sum over rows, and an entity a cube takes from a neighbour holds
    // the number its grid position got from the cube that created it. The
    // first cube to hold an entity creates it, so the creating cube sets what
    // the reference sets on an entity's first visit.
    struct NewSet { int nf = 0, ne = 0, np = 0; int f[6], e[12], p[8]; uint8_t fnew = 0; };
    NewSet nset[8];  // by neighbour mask: 1 bottom, 2 left, 4 back
    for (int m = 0; m < 8; ++m) {
        auto reused = [m](const std::vector<Src>& r) {
            for (const Src& x : r)
                if (m >> x.nb & 1) return true;
            return false;
        };
        NewSet& t = nset[m];
        for (int f = 0; f < 6; ++f)
            if (!reused(kFaceReuse[f])) { t.f[t.nf++] = f; t.fnew |= (uint8_t)(1u << f); }
        for (int e = 0; e < 12; ++e)
            if (!reused(kEdgeReuse[e])) t.e[t.ne++] = e;
        for (int q = 0; q < 8; ++q)
            if (!reused(kNodeReuse[q])) t.p[t.np++] = q;
    }
    std::vector<uint8_t> has_back(Nx);
    int64_t n_back = 0;  // cubes of a row with a back neighbour
    for (uint32_t i = 0; i < Nx; ++i) n_back += has_back[i] = xs[i] > xmin;
    auto row_mask = [&](uint32_t j, uint32_t k) { return (zs[k] > zmin ? 1 : 0) | (ys[j] > ymin ? 2 : 0); };
    std::vector<int64_t> rowF(nrow + 1, 0), rowE(nrow + 1, 0), rowP(nrow + 1, 0);
    for (int64_t r = 0; r < nrow; ++r) {
        const int m = row_mask((uint32_t)(r % Ny), (uint32_t)(r / Ny));
        const NewSet &t0 = nset[m], &t1 = nset[m | 4];
        rowF[r + 1] = rowF[r] + ((int64_t)Nx - n_back) * t0.nf + n_back * t1.nf;
        rowE[r + 1] = rowE[r] + ((int64_t)Nx - n_back) * t0.ne + n_back * t1.ne;
        rowP[r + 1] = rowP[r] + ((int64_t)Nx - n_back) * t0.np + n_back * t1.np;
    }
    if ((uint64_t)rowF[nrow] != totF || (uint64_t)rowE[nrow] != totE || (uint64_t)rowP[nrow] != totP)
        throw Error("cartesian mesher: entity count mismatch");
    auto cube_face_pos = [&](int f, int64_t i, int64_t j, int64_t k) {
        const int a = kFaceAxis[f], o = f >= 3 ? 1 : 0;
        return face_pos(a, i + (a == 0 ? o : 0), j + (a == 1 ? o : 0), k + (a == 2 ? o : 0));
    };
    auto cube_edge_pos = [&](int e, int64_t i, int64_t j, int64_t k) {
        const int q0 = kEdgeNodes[e][0];  // the edge's lower node
        return edge_pos(kEdgeAxis[e], i + (q0 & 1), j + (q0 >> 1 & 1), k + (q0 >> 2 & 1));
    };
    auto cube_node_pos = [&](int q, int64_t i, int64_t j, int64_t k) {
        return ((k + (q >> 2 & 1)) * (Ny + 1) + (j + (q >> 1 & 1))) * (Nx + 1) + (i + (q & 1));
    };
    // number of the entity at each grid position
    ParTable<int32_t> f_at(totF, 0), e_at(totE, 0), p_at(totP, 0);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < nrow; ++r) {
        const uint32_t j = (uint32_t)(r % Ny), k = (uint32_t)(r / Ny);
        const int m = row_mask(j, k);
        int64_t nf = rowF[r], ne = rowE[r], np = rowP[r];
        for (uint32_t i = 0; i < Nx; ++i) {
            const NewSet& t = nset[m | (has_back[i] ? 4 : 0)];
            for (int u = 0; u < t.nf; ++u) f_at[cube_face_pos(t.f[u], i, j, k)] = (int32_t)nf++;
            for (int u = 0; u < t.ne; ++u) e_at[cube_edge_pos(t.e[u], i, j, k)] = (int32_t)ne++;
            const Vec3 pp(xs[i], ys[j], zs[k]);
            for (int u = 0; u < t.np; ++u) {
                const int q = t.p[u];
                const int64_t id = np++;
                p_at[cube_node_pos(q, i, j, k)] = (int32_t)id;
                Vec3 p = pp;
                if (q & 4) p = p + Vec3(0, 0, Lz);
                if (q & 1) p = p + Vec3(Lx, 0, 0);
                if (q & 2) p = p + Vec3(0, Ly, 0);
                b.pts[id * 3] = p.x; b.pts[id * 3 + 1] = p.y; b.pts[id * 3 + 2] = p.z;
            }
        }
    }
    // cube tables, incidences and what the creating cube sets
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < nrow; ++r) {
        const uint32_t j = (uint32_t)(r % Ny), k = (uint32_t)(r / Ny);
        const int m = row_mask(j, k);
        for (uint32_t i = 0; i < Nx; ++i) {
            const NewSet& t = nset[m | (has_back[i] ? 4 : 0)];
            const size_t c = (size_t)r * Nx + i;
            int32_t dcube[6], ecube[12], pcube[8];
            for (int f = 0; f < 6; ++f) D[c * 6 + f] = dcube[f] = f_at[cube_face_pos(f, i, j, k)];
            for (int e = 0; e < 12; ++e) E_cluster[c * 12 + e] = ecube[e] = e_at[cube_edge_pos(e, i, j, k)];
            for (int q = 0; q < 8; ++q) P_cluster[c * 8 + q] = pcube[q] = p_at[cube_node_pos(q, i, j, k)];
            // the lower cube of a face comes first in the walk: a face taken
            // from the neighbour below / left / behind is its second visit
            for (int f = 0; f < 6; ++f) {
                const size_t slot = (size_t)dcube[f] * 2 + (t.fnew >> f & 1 ? 0 : 1);
                Dt[slot] = (int32_t)c;
                Dt_sgn[slot] = f < 3 ? -1 : 1;
            }
            for (int u = 0; u < t.ne; ++u) {
                const int e = t.e[u];
                const int32_t eid = ecube[e];
                b.edge_sid[eid] = cube_edge_pos(e, i, j, k);
                G[(size_t)eid * 2] = pcube[kEdgeNodes[e][0]];
                G[(size_t)eid * 2 + 1] = pcube[kEdgeNodes[e][1]];
                edge_len[eid] = len[kEdgeAxis[e]];
                b.dual_curl[eid] = (int8_t)kDualCurl[kEdgeAxis[e]];
            }
            for (int u = 0; u < t.nf; ++u) {
                const int f = t.f[u];
                const int32_t fid = dcube[f];
                b.face_sid[fid] = cube_face_pos(f, i, j, k);
                for (int q = 0; q < 4; ++q) {
                    const int32_t eid = ecube[kFaceEdges[f][q]];
                    b.c_edge[(size_t)fid * 4 + q] = eid;
                    const uint8_t n = __atomic_fetch_add(&ct_n[eid], (uint8_t)1, __ATOMIC_RELAXED);
                    ct_face[(size_t)eid * 4 + n] = fid;
                    ct_sgn[(size_t)eid * 4 + n] = (int8_t)kFaceSigns[f][q];
                }
                b.curl[fid] = (int8_t)kFaceCurl[f];
                face_area[fid] = area[kFaceAxis[f]];
                boundary_face[fid] = (uint32_t)kFaceLabel[f];
            }
        }
    }
    // faces: cubes, labels and material averages (D:5244-5330), cube by cube
    // in walk order like the reference's sums
#pragma omp parallel for schedule(static)
    for (int64_t fid = 0; fid < (int64_t)totF; ++fid) {
        const int32_t c0 = Dt[(size_t)fid * 2], c1 = Dt[(size_t)fid * 2 + 1];
        Dt_n[fid] = c1 >= 0 ? 2 : 1;
        tbc[fid] = vol_material[c0] != 0;
        if (c1 >= 0) {
            boundary_face[fid] = matad_at(vol_material[c0], vol_material[c1]);
            tbc[fid] = tbc[fid] || vol_material[c1] != 0;
        }
        const int64_t fp = b.face_sid[fid];
        const int a = fp < nFx ? 0 : fp < nFx + nFy ? 1 : 2;  // normal axis
        for (int32_t c : {c0, c1}) {
            if (c < 0) continue;
            const Material& mt = *solid_mat[vol_solid[c]];
            const double mu = mt.mu(a, a), ch = mt.chi(a, a);
            if (mu != 0) avg_ni[fid] += len[a] / 2 / mu;
            if (ch != 0) { avg_mag_sigma[fid] += len[a] / 2 / ch; mag_lossy[fid]++; }
        }
    }
    // edges: faces in creation (= number) order, and material averages over
    // the up to four cubes around the edge in walk order
#pragma omp parallel for schedule(static)
    for (int64_t eid = 0; eid < (int64_t)totE; ++eid) {
        int32_t* cf = &ct_face[(size_t)eid * 4];
        int8_t* cs = &ct_sgn[(size_t)eid * 4];
        for (int x = 1; x < ct_n[eid]; ++x)
            for (int y = x; y > 0 && cf[y - 1] > cf[y]; --y) { std::swap(cf[y - 1], cf[y]); std::swap(cs[y - 1], cs[y]); }
        int64_t rem = b.edge_sid[eid];
        const int a = rem < nEx ? 0 : rem < nEx + nEy ? 1 : 2;
        rem -= a == 0 ? 0 : a == 1 ? nEx : nEx + nEy;
        const int64_t ni = Nx + (a != 0), nj = Ny + (a != 1);  // positions per line and per plane row
        const int64_t i = rem % ni, j = rem / ni % nj, k = rem / ni / nj;
        // the two axes across the edge, lower one first: the cubes at
        // (-1,-1), (0,-1), (-1,0), (0,0) in them come in walk order
        const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
        for (int s = 0; s < 4; ++s) {
            int64_t q[3] = {i, j, k};
            q[u] -= (s & 1) ? 0 : 1;
            q[v] -= (s & 2) ? 0 : 1;
            if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= Nx || q[1] >= Ny || q[2] >= Nz) continue;
            const size_t c = ((size_t)q[2] * Ny + q[1]) * Nx + q[0];
            const Material& mt = *solid_mat[vol_solid[c]];
            double ep = mt.epsilon(a, a);
            if (mt.eps_random) ep = EPSILON0 * (mt.eps_lo + (mt.eps_hi - mt.eps_lo) * cell_uniform(mt.eps_seed, c));
            const double si = mt.sigma(a, a);
            if (ep != 0) avg_eps[eid] += da[a] * ep;
            if (si != 0) { avg_sigma[eid] += da[a] * si; ele_lossy[eid]++; }
        }
    }
    const uint32_t nv = (uint32_t)totV, np = (uint32_t)totP;  // cubes and nodes
    b.n_cubes = (int32_t)nv;

    clk.mark("mesher (cubes, numbering, averages)");
    // ---- Ct_vec as CSR
    b.ct_ptr.assign(totE + 1, 0);
    for (uint64_t e = 0; e < totE; ++e) b.ct_ptr[e + 1] = b.ct_ptr[e] + ct_n[e];
    b.ct_face.resize(b.ct_ptr[totE]);
    b.ct_sgn.resize(b.ct_ptr[totE]);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < (int64_t)totE; ++e)
        for (int q = 0; q < ct_n[e]; ++q) {
            b.ct_face[b.ct_ptr[e] + q] = ct_face[e * 4 + q];
            b.ct_sgn[b.ct_ptr[e] + q] = ct_sgn[e * 4 + q];
        }
    ct_face = ParTable<int32_t>();

    clk.mark("Ct");
    // ---- edge classes (D:5382-5471). An edge none of whose faces lies on
    // the boundary is common; the others (a thin shell) run the reference's
    // per-face classification in edge order. The sources of an edge are kept
    // only for the edges that have some (a sparse map).
    std::map<uint32_t, std::vector<uint32_t>> edge_src;
    std::vector<uint32_t> edge_bcs(totE, 0);
    auto bc_of = [&](uint32_t label) -> const BoundaryCondition& {
        uint32_t l = label < m.bnd_to_bc.size() ? m.bnd_to_bc[label] : 0;
        return bcs[l];
    };
    std::vector<uint8_t> common(totE, 1), shell(totE, 0);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < (int64_t)totE; ++e)
        for (int64_t q = b.ct_ptr[e]; q < b.ct_ptr[e + 1]; ++q)
            if (boundary_face[b.ct_face[q]] > 0) { shell[e] = 1; break; }
    for (uint32_t e = 0; e < totE; ++e) {
        if (!shell[e]) continue;
        uint32_t in_b = 0, brk = 0;
        for (int64_t q = b.ct_ptr[e]; q < b.ct_ptr[e + 1]; ++q) {
            uint32_t ff = (uint32_t)b.ct_face[q];
            uint32_t bid = (uint8_t)boundary_face[ff];
            if (boundary_face[ff] > 0) {
                if (bc_of(bid).type == "pec") {
                    ++in_b;
                    edge_bcs[e] = bid < m.bnd_to_bc.size() ? m.bnd_to_bc[bid] : 0;
                    edge_src.erase(e);
                    b.bc_edges.push_back((int32_t)e);
                    if (brk == 1) b.src_edges.pop_back();
                    else if (brk == 2) b.uncommon_edges.pop_back();
                    break;
                } else {
                    for (uint32_t sl : sim.sources) {
                        const SourceDef& src = sources[sl];
                        if (std::find(src.surface_label.begin(), src.surface_label.end(), bid) == src.surface_label.end())
                            continue;
                        ++in_b;
                        if (src.st == "e") {
                            if (brk == 2) break;
                            std::vector<uint32_t>& es = edge_src[e];
                            if (std::find(es.begin(), es.end(), sl) == es.end()) {
                                es.push_back(sl);
                                brk = 1;
                                if (b.src_edges.empty() || b.src_edges.back() != (int32_t)e) b.src_edges.push_back((int32_t)e);
                            }
                        } else if (src.st == "h") {
                            if (brk == 1) break;
                            std::vector<uint32_t>& es = edge_src[e];
                            if (std::find(es.begin(), es.end(), sl) == es.end()) {
                                es.push_back(sl);
                                brk = 2;
                                if (b.uncommon_edges.empty() || b.uncommon_edges.back() != (int32_t)e)
                                    b.uncommon_edges.push_back((int32_t)e);
                            }
                        }
                    }
                }
            }
        }
        common[e] = in_b == 0;
    }
    {
        // common edges in edge order: counts per chunk, then a parallel fill
        constexpr int64_t kChunk = 1 << 20;
        const int64_t nch = ((int64_t)totE + kChunk - 1) / kChunk;
        std::vector<int64_t> off(nch + 1, 0);
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < nch; ++c) {
            int64_t k = 0;
            for (int64_t e = c * kChunk; e < std::min<int64_t>(totE, (c + 1) * kChunk); ++e) k += common[e];
            off[c + 1] = k;
        }
        for (int64_t c = 0; c < nch; ++c) off[c + 1] += off[c];
        b.common_edges.resize(off[nch]);
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < nch; ++c) {
            int64_t k = off[c];
            for (int64_t e = c * kChunk; e < std::min<int64_t>(totE, (c + 1) * kChunk); ++e)
                if (common[e]) b.common_edges[k++] = (int32_t)e;
        }
    }
    for (int32_t e : b.common_edges)
        if (b.ct_ptr[e + 1] - b.ct_ptr[e] != 4)
            throw Error("common edge without 4 faces: every outer wall must be pec or a source surface (D:1872)");
    // ComputeEdgeBC: BCs[BndToBC[edge_bcs[e]]].GetValue() (D:3899-3905)
    for (int32_t e : b.bc_edges) b.bc_value.push_back(bc_of(edge_bcs[e]).val);
    for (uint32_t f = 0; f < totF; ++f)
        if (tbc[f]) b.tbc_faces.push_back((int32_t)f);

    clk.mark("edge classes, bc");
    // ---- time step: power iteration on the FIT curl-curl (D:7036-7119),
    // Hv Ct Nv C with Nv = avg_ni / area per face and Hv = length / avg_eps
    // per edge. It runs in grid order: each face's four edges and each
    // edge's (up to four) faces are found from the grid position, with the
    // signs C and Ct give them (read off c_edge / curl and ct_sgn), so the
    // operator is the same entry by entry and only the storage order differs;
    // the start vector is the same pseudo-random vector, placed at the grid
    // positions. Power iteration approaches lambda_max from below: pad by 2%.
    {
        const int64_t nEz = (int64_t)(Nx + 1) * (Ny + 1) * Nz, nFz = (int64_t)Nx * Ny * (Nz + 1);
        (void)nEz; (void)nFz;
        // grid position -> (axis, i, j, k)
        auto dec_edge = [&](int64_t g, int64_t* q) -> int {
            if (g < nEx) { q[0] = g % Nx; q[1] = g / Nx % (Ny + 1); q[2] = g / Nx / (Ny + 1); return 0; }
            if (g < nEx + nEy) { g -= nEx; q[0] = g % (Nx + 1); q[1] = g / (Nx + 1) % Ny; q[2] = g / (Nx + 1) / Ny; return 1; }
            g -= nEx + nEy; q[0] = g % (Nx + 1); q[1] = g / (Nx + 1) % (Ny + 1); q[2] = g / (Nx + 1) / (Ny + 1); return 2;
        };
        auto dec_face = [&](int64_t g, int64_t* q) -> int {
            if (g < nFx) { q[0] = g % (Nx + 1); q[1] = g / (Nx + 1) % Ny; q[2] = g / (Nx + 1) / Ny; return 0; }
            if (g < nFx + nFy) { g -= nFx; q[0] = g % Nx; q[1] = g / Nx % (Ny + 1); q[2] = g / Nx / (Ny + 1); return 1; }
            g -= nFx + nFy; q[0] = g % Nx; q[1] = g / Nx % Ny; q[2] = g / Nx / Ny; return 2;
        };
        // the four edges of a face: x-face y(i,j,k) z(i,j+1,k) y(i,j,k+1) z(i,j,k); y-face z(i,j,k)
        // x(i,j,k+1) z(i+1,j,k) x(i,j,k); z-face x(i,j,k) y(i+1,j,k) x(i,j+1,k) y(i,j,k)
        auto face_edges = [&](int a, int64_t i, int64_t j, int64_t k, int32_t* e4) {
            if (a == 0) { e4[0] = edge_pos(1, i, j, k); e4[1] = edge_pos(2, i, j + 1, k); e4[2] = edge_pos(1, i, j, k + 1); e4[3] = edge_pos(2, i, j, k); }
            else if (a == 1) { e4[0] = edge_pos(2, i, j, k); e4[1] = edge_pos(0, i, j, k + 1); e4[2] = edge_pos(2, i + 1, j, k); e4[3] = edge_pos(0, i, j, k); }
            else { e4[0] = edge_pos(0, i, j, k); e4[1] = edge_pos(1, i + 1, j, k); e4[2] = edge_pos(0, i, j + 1, k); e4[3] = edge_pos(1, i, j, k); }
        };
        // the faces of an edge (-1: outside the grid): x-edge y(i,j,k) y(i,j,k-1) z(i,j,k)
        // z(i,j-1,k); y-edge z(i,j,k) z(i-1,j,k) x(i,j,k) x(i,j,k-1); z-edge x(i,j,k)
        // x(i,j-1,k) y(i,j,k) y(i-1,j,k)
        auto edge_faces = [&](int a, int64_t i, int64_t j, int64_t k, int32_t* f4) {
            if (a == 0) {
                f4[0] = k < Nz ? face_pos(1, i, j, k) : -1; f4[1] = k > 0 ? face_pos(1, i, j, k - 1) : -1;
                f4[2] = j < Ny ? face_pos(2, i, j, k) : -1; f4[3] = j > 0 ? face_pos(2, i, j - 1, k) : -1;
            } else if (a == 1) {
                f4[0] = i < Nx ? face_pos(2, i, j, k) : -1; f4[1] = i > 0 ? face_pos(2, i - 1, j, k) : -1;
                f4[2] = k < Nz ? face_pos(0, i, j, k) : -1; f4[3] = k > 0 ? face_pos(0, i, j, k - 1) : -1;
            } else {
                f4[0] = j < Ny ? face_pos(0, i, j, k) : -1; f4[1] = j > 0 ? face_pos(0, i, j - 1, k) : -1;
                f4[2] = i < Nx ? face_pos(1, i, j, k) : -1; f4[3] = i > 0 ? face_pos(1, i - 1, j, k) : -1;
            }
        };
        // per grid face / edge: two bits per slot, 1 = +1, 2 = -1, 0 = no entry
        std::vector<uint8_t> fsg(totF, 0), esg(totE, 0);
        std::vector<double> Nv(totF), Hv(totE);
        bool ok = true;
#pragma omp parallel for schedule(static) reduction(&& : ok)
        for (int64_t f = 0; f < (int64_t)totF; ++f) {
            const int32_t g = b.face_sid[f];
            int64_t q[3];
            const int a = dec_face(g, q);
            int32_t e4[4];
            face_edges(a, q[0], q[1], q[2], e4);
            uint8_t code = 0;
            for (int t = 0; t < 4; ++t) {
                const int32_t ge = b.edge_sid[b.c_edge[(size_t)f * 4 + t]];
                const int sgn = b.curl[f] * (t & 1 ? -1 : 1);
                int slot = -1;
                for (int u = 0; u < 4; ++u) if (e4[u] == ge) slot = u;
                if (slot < 0) ok = false;
                else code |= (uint8_t)((sgn > 0 ? 1 : 2) << (2 * slot));
            }
            fsg[g] = code;
            Nv[g] = avg_ni[f] / face_area[f];
        }
#pragma omp parallel for schedule(static) reduction(&& : ok)
        for (int64_t e = 0; e < (int64_t)totE; ++e) {
            const int32_t g = b.edge_sid[e];
            int64_t q[3];
            const int a = dec_edge(g, q);
            int32_t f4[4];
            edge_faces(a, q[0], q[1], q[2], f4);
            uint8_t code = 0;
            for (int64_t t = b.ct_ptr[e]; t < b.ct_ptr[e + 1]; ++t) {
                const int32_t gf = b.face_sid[b.ct_face[t]];
                int slot = -1;
                for (int u = 0; u < 4; ++u) if (f4[u] == gf) slot = u;
                if (slot < 0 || (code >> (2 * slot) & 3)) ok = false;
                else code |= (uint8_t)((b.ct_sgn[t] > 0 ? 1 : 2) << (2 * slot));
            }
            esg[g] = code;
            Hv[g] = edge_len[e] / avg_eps[e];
        }
        if (!ok) throw Error("internal: FIT incidence does not match the grid");
        // row bases of the grid arrays (entity (i, j, k) of a row at base + i)
        const int64_t NX = Nx, NY = Ny, NZ = Nz;
        auto EX = [&](int64_t j, int64_t k) { return (k * (NY + 1) + j) * NX; };
        auto EY = [&](int64_t j, int64_t k) { return nEx + (k * NY + j) * (NX + 1); };
        auto EZ = [&](int64_t j, int64_t k) { return nEx + nEy + (k * (NY + 1) + j) * (NX + 1); };
        auto FX = [&](int64_t j, int64_t k) { return (k * NY + j) * (NX + 1); };
        auto FY = [&](int64_t j, int64_t k) { return nFx + (k * (NY + 1) + j) * NX; };
        auto FZ = [&](int64_t j, int64_t k) { return nFx + nFy + (k * NY + j) * NX; };
        // one row of entities at g0 .. g0 + n: out[g] = w[g] * sum over the
        // four slots u with a sign code of (+-) in[base[u] + i]
        auto row = [](const uint8_t* code, const double* w, const double* in, double* out, int64_t g0, int64_t n,
                      const int64_t* base) {
            for (int64_t i = 0; i < n; ++i) {
                const int c = code[g0 + i];
                double s = 0;
                for (int u = 0; u < 4; ++u) {
                    const int cu = c >> (2 * u) & 3;
                    if (cu) s += (cu == 1 ? 1.0 : -1.0) * in[base[u] + i];
                }
                out[g0 + i] = w[g0 + i] * s;
            }
        };
        std::vector<double> Fw(totF);
        auto apply = [&](const std::vector<double>& X, std::vector<double>& Y) {
            const double* x = X.data();
            double* fw = Fw.data();
#pragma omp parallel for schedule(static)
            for (int64_t r = 0; r < NZ * NY + (NZ * (NY + 1)) + (NZ + 1) * NY; ++r) {
                int64_t base[4];
                if (r < NZ * NY) {  // x faces, rows (j < Ny, k < Nz) of Nx + 1
                    const int64_t j = r % NY, k = r / NY;
                    base[0] = EY(j, k); base[1] = EZ(j + 1, k); base[2] = EY(j, k + 1); base[3] = EZ(j, k);
                    row(fsg.data(), Nv.data(), x, fw, FX(j, k), NX + 1, base);
                } else if (r < NZ * NY + NZ * (NY + 1)) {  // y faces, rows (j <= Ny, k < Nz) of Nx
                    const int64_t q = r - NZ * NY, j = q % (NY + 1), k = q / (NY + 1);
                    base[0] = EZ(j, k); base[1] = EX(j, k + 1); base[2] = EZ(j, k) + 1; base[3] = EX(j, k);
                    row(fsg.data(), Nv.data(), x, fw, FY(j, k), NX, base);
                } else {  // z faces, rows (j < Ny, k <= Nz) of Nx
                    const int64_t q = r - NZ * NY - NZ * (NY + 1), j = q % NY, k = q / NY;
                    base[0] = EX(j, k); base[1] = EY(j, k) + 1; base[2] = EX(j + 1, k); base[3] = EY(j, k);
                    row(fsg.data(), Nv.data(), x, fw, FZ(j, k), NX, base);
                }
            }
            double* y = Y.data();
#pragma omp parallel for schedule(static)
            for (int64_t r = 0; r < (NZ + 1) * (NY + 1) + (NZ + 1) * NY + NZ * (NY + 1); ++r) {
                int64_t base[4];
                if (r < (NZ + 1) * (NY + 1)) {  // x edges, rows (j <= Ny, k <= Nz) of Nx
                    const int64_t j = r % (NY + 1), k = r / (NY + 1);
                    base[0] = FY(j, k); base[1] = FY(j, k - 1); base[2] = FZ(j, k); base[3] = FZ(j - 1, k);
                    row(esg.data(), Hv.data(), fw, y, EX(j, k), NX, base);
                } else if (r < (NZ + 1) * (NY + 1) + (NZ + 1) * NY) {  // y edges, rows (j < Ny, k <= Nz) of Nx + 1
                    const int64_t q = r - (NZ + 1) * (NY + 1), j = q % NY, k = q / NY;
                    base[0] = FZ(j, k); base[1] = FZ(j, k) - 1; base[2] = FX(j, k); base[3] = FX(j, k - 1);
                    row(esg.data(), Hv.data(), fw, y, EY(j, k), NX + 1, base);
                } else {  // z edges, rows (j <= Ny, k < Nz) of Nx + 1
                    const int64_t q = r - (NZ + 1) * (NY + 1) - (NZ + 1) * NY, j = q % (NY + 1), k = q / (NY + 1);
                    base[0] = FX(j, k); base[1] = FX(j - 1, k); base[2] = FY(j, k); base[3] = FY(j, k) - 1;
                    row(esg.data(), Hv.data(), fw, y, EZ(j, k), NX + 1, base);
                }
            }
        };
        std::vector<double> x0(totE);
        {
            uint64_t st = 0x9E3779B97F4A7C15ull;
            for (uint64_t e = 0; e < totE; ++e) {
                st ^= st << 13; st ^= st >> 7; st ^= st << 17;
                x0[b.edge_sid[e]] = (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
            }
        }
        double lam = power_lambda_max(totE, apply, 60, &x0) * 1.02;
        b.stable_dt = lam > 0 ? 2.0 / std::sqrt(lam) : 0;
    }
    double t_step = sim.courant * b.stable_dt;
    if (sim.step != 0) {
        if (sim.step > t_step) b.possibly_unstable = true;
        else t_step = sim.step;
    }
    b.prob.dt = t_step;

    clk.mark("dt estimate");
    // ---- coefficients (D:5487-5649). PML factors are 1 without PML.
    b.M_nu.resize(totF); b.M_mu.resize(totF); b.fpml.assign(totF, 1.0); b.Nu_vec.assign(totF, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < (int64_t)totF; ++f) {
        if (avg_ni[f] != 0) {
            b.Nu_vec[f] = avg_ni[f] / face_area[f];
            if (mag_lossy[f]) {
                b.M_nu[f] = 1 / (face_area[f] * (1 / avg_ni[f] + 0.5 * t_step / avg_mag_sigma[f]));
                b.M_mu[f] = face_area[f] * (1 / avg_ni[f] - 0.5 * t_step / avg_mag_sigma[f]);
            } else {
                b.M_nu[f] = avg_ni[f] / face_area[f];
                b.M_mu[f] = face_area[f] / avg_ni[f];
            }
        } else { b.M_nu[f] = 0; b.M_mu[f] = 0; }
    }
    for (auto& kv : edge_src) {  // source consistency (D:5573-5586)
        std::vector<uint32_t>& es = kv.second;
        if (es.empty()) continue;
        std::vector<uint32_t> keep;
        const SourceDef& s0 = sources[es.front()];
        uint32_t surf0 = s0.surface_label.empty() ? 0 : s0.surface_label.front();
        for (uint32_t k : es) {
            const SourceDef& sk = sources[k];
            uint32_t sf = sk.surface_label.empty() ? 0 : sk.surface_label.front();
            if (sk.st == s0.st && sf == surf0) keep.push_back(k);
        }
        es = keep;
    }
    b.M_h.resize(totE); b.M_q.resize(totE); b.M_e.resize(totE); b.upml.assign(totE, 1.0);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < (int64_t)totE; ++e) {
        if (avg_eps[e] != 0) {
            if (ele_lossy[e]) {
                b.M_h[e] = edge_len[e] / (avg_eps[e] + 0.5 * t_step * avg_sigma[e]);
                b.M_q[e] = (avg_eps[e] - 0.5 * t_step * avg_sigma[e]) / edge_len[e];
            } else {
                b.M_h[e] = edge_len[e] / avg_eps[e];
                b.M_q[e] = avg_eps[e] / edge_len[e];
            }
            b.M_e[e] = avg_eps[e] / edge_len[e];
        } else { b.M_h[e] = b.M_q[e] = b.M_e[e] = 0; }
    }
    b.boundary_face.assign(boundary_face.begin(), boundary_face.end());

    clk.mark("coefficients");
    // ---- geometry helpers
    auto P = [&](int32_t n) { return Vec3(b.pts[(size_t)n * 3], b.pts[(size_t)n * 3 + 1], b.pts[(size_t)n * 3 + 2]); };
    auto edge_bar = [&](uint32_t e) {
        Vec3 bc(0, 0, 0);
        bc = bc + P(G[(size_t)e * 2]);
        bc = bc + P(G[(size_t)e * 2 + 1]);
        return bc / 2.0;
    };
    auto face_bar = [&](uint32_t f) {
        Vec3 bc(0, 0, 0);
        for (int q = 0; q < 4; ++q) {
            int32_t e = b.c_edge[(size_t)f * 4 + q];
            bc = bc + 0.5 * P(G[(size_t)e * 2]);
            bc = bc + 0.5 * P(G[(size_t)e * 2 + 1]);
        }
        return bc / 4.0;
    };
    auto vol_bar = [&](uint32_t v) {
        Vec3 bc(0, 0, 0);
        for (int q = 0; q < 8; ++q) bc = bc + P(P_cluster[(size_t)v * 8 + q]);
        return bc / 8.0;
    };
    // PML factors (D:5519-5557 faces, D:5609-5647 edges): beyond the inner
    // box on axis a, cf *= exp(-(d / width_a)^3 * 4 / 2 / 150 / pi / L_a) with
    // d the distance past the box and width_a = true_max - max, evaluated
    // in long double like the reference (its PI is long double)
    if (pml_present) {
        const long double PI_L = 3.141592653589793238L;
        const double inner[3] = {xmax, ymax, zmax};
        auto pml = [&](const Vec3& c) {
            double cf = 1;
            for (int a = 0; a < 3; ++a)
                if (Npml[a] > 0 && c[a] > inner[a]) {
                    const double d = std::fabs(inner[a] - c[a]) / std::fabs(inner[a] - true_max[a]);
                    cf *= std::exp(std::pow(d, 3) * 4 / 2 / 150 / PI_L / len[a] * -1.0L);
                }
            return cf;
        };
        for (uint32_t f = 0; f < totF; ++f) b.fpml[f] = pml(face_bar(f));
        for (uint32_t e = 0; e < totE; ++e) b.upml[e] = pml(edge_bar(e));
    }

    clk.mark("geometry helpers");
    // ---- probes (FindFitProbe, D:3590-3643) and j dipoles (D:447-492)
    set_output(m, sim, b);
    uint32_t root = 0;
    auto find_fit_probe = [&](const Vec3& pv) -> uint32_t {
        const double tol = 1e-12;
        std::vector<uint8_t> colour(nv, 0);
        std::vector<uint32_t> queue;
        queue.push_back(root);
        colour[0]++;
        for (size_t k = 0; k < queue.size(); ++k) {
            uint32_t v = queue[k];
            Vec3 diff = pv - P(P_cluster[(size_t)v * 8]);
            bool c1 = diff.x >= -tol && diff.x < Lx + tol;
            bool c2 = diff.y >= -tol && diff.y < Ly + tol;
            bool c3 = diff.z >= -tol && diff.z < Lz + tol;
            if (c1 && c2 && c3) { root = v; return v; }
            for (int f = 0; f < 6; ++f) {
                int32_t fid = D[(size_t)v * 6 + f];
                uint32_t cn = (uint32_t)Dt[(size_t)fid * 2];
                if (cn == v) cn = (uint32_t)Dt[(size_t)fid * 2 + Dt_n[fid] - 1];
                if (colour[cn] == 0) { queue.push_back(cn); colour[cn]++; }
            }
            colour[v]++;
        }
        return nv;
    };
    if (b.out_mode == 1 || b.out_mode == 2) {
        const OutputDef& o = m.outputs.at(sim.output);
        const double vol = Lx * Ly * Lz;
        const Vec3 dual_area[3] = {Vec3(Ly * Lz / 4, 0, 0), Vec3(0, Lx * Lz / 4, 0), Vec3(0, 0, Lx * Ly / 4)};
        const Vec3 inc[3] = {Vec3(0.5 * Lx, 0, 0), Vec3(0, 0.5 * Ly, 0), Vec3(0, 0, 0.5 * Lz)};
        for (const Vec3& pv : o.probes) {
            uint32_t c = find_fit_probe(pv);
            if (c >= nv) continue;  // out of mesh: ignored with a warning
            tf_probe pr{};
            tf_probe_geom g{};
            pr.ne = 12;
            pr.nh = 6;
            for (int q = 0; q < 12; ++q) {
                pr.e_idx[q] = E_cluster[(size_t)c * 12 + q];
                Vec3 w = dual_area[kEdgeAxis[q]] / vol;
                pr.e_w[q][0] = w.x; pr.e_w[q][1] = w.y; pr.e_w[q][2] = w.z;
                g.edges[q] = pr.e_idx[q];
            }
            Mat3 mui = materials[vol_material[c]].mu.inverse();
            for (int q = 0; q < 6; ++q) {
                pr.h_idx[q] = D[(size_t)c * 6 + q];
                Vec3 w = (mui * inc[kFaceAxis[q]]) / vol;
                pr.h_w[q][0] = w.x; pr.h_w[q][1] = w.y; pr.h_w[q][2] = w.z;
                g.faces[q] = pr.h_idx[q];
            }
            g.elem = (int32_t)c;
            g.point[0] = pv.x; g.point[1] = pv.y; g.point[2] = pv.z;
            g.L[0] = Lx; g.L[1] = Ly; g.L[2] = Lz;
            std::memcpy(g.mu, materials[vol_material[c]].mu.m, sizeof(g.mu));
            b.probes.push_back(pr);
            b.probe_geom.push_back(g);
            b.probe_points.push_back(pv.x); b.probe_points.push_back(pv.y); b.probe_points.push_back(pv.z);
        }
        for (uint32_t sl : sim.sources) {
            SourceDef& s = sources[sl];
            if (s.st != "j") continue;
            Vec3 loc(s.location[0], s.location[1], s.location[2]);
            uint32_t c = find_fit_probe(loc);
            if (c >= nv) continue;
            int32_t best = E_cluster[(size_t)c * 12];
            double bd = norm(edge_bar(best) - loc);
            for (int q = 1; q < 12; ++q) {
                int32_t e = E_cluster[(size_t)c * 12 + q];
                double d = norm(edge_bar(e) - loc);
                if (d < bd) { bd = d; best = e; }
            }
            if (edge_src[best].empty()) b.src_edges.push_back(best);
            edge_src[best].push_back(sl);
        }
    }

    clk.mark("probes");
    // ---- sources: per-edge weights. h sources on uncommon edges use the
    // boundary dual edge vector (D:4148-4199); e/j sources the edge vector.
    std::map<uint32_t, int32_t> src_index;
    for (uint32_t sl : sim.sources) {
        src_index[sl] = (int32_t)b.sources.size();
        b.sources.push_back(make_source(sources[sl]));
        b.src_desc.push_back(describe_source(sources[sl]));
    }
    auto bnd_dual = [&](uint32_t e) {
        Vec3 eb = edge_bar(e), v(0, 0, 0), vb(0, 0, 0), backup(0, 0, 0);
        int nfb = 0, ni = 0;
        for (int64_t q = b.ct_ptr[e]; q < b.ct_ptr[e + 1]; ++q) {
            uint32_t ff = (uint32_t)b.ct_face[q];
            if (boundary_face[ff] > 0) {
                ++nfb;
                backup = vol_bar((uint32_t)Dt[(size_t)ff * 2]) - eb;
                Vec3 fb = face_bar(ff);
                if (nfb == 1) v = fb;
                else if (nfb == 2) v = fb - v;
            } else {
                vb = face_bar(ff) - eb;
                ++ni;
            }
        }
        if (!ni) vb = backup;
        Vec3 ev = P(G[(size_t)e * 2 + 1]) - P(G[(size_t)e * 2]);
        if (dot(cross(ev, v), vb) < 0) v = -v;
        return v;
    };
    auto add_edge = [&](uint32_t e) {
        auto es_it = edge_src.find(e);
        if (es_it == edge_src.end() || es_it->second.empty()) return;
        const std::vector<uint32_t>& es = es_it->second;
        tf_source_edge d{};
        d.edge = (int32_t)e;
        const SourceDef& s0 = sources[es.front()];
        d.field = make_source(s0).field;  // only e, h (D:5431-5455) and j dipoles reach edge_src here
        if (es.size() > 4)
            throw Error("more than 4 sources on one edge (edge " + std::to_string(e) + ")");
        Vec3 eb = edge_bar(e);
        Vec3 ev = P(G[(size_t)e * 2 + 1]) - P(G[(size_t)e * 2]);
        Vec3 vec = d.field == TF_FIELD_H ? bnd_dual(e) : ev;
        d.bary[0] = eb.x; d.bary[1] = eb.y; d.bary[2] = eb.z;
        d.vec[0] = vec.x; d.vec[1] = vec.y; d.vec[2] = vec.z;
        for (uint32_t sl : es) {
            const SourceDef& s = sources[sl];
            d.src[d.n_src++] = src_index[sl];
            if (make_source(s).field != d.field) continue;  // the loop only evaluates sources of the edge's kind
            Vec3 sp = SourceEval::space(s, eb);
            double w = dot(sp, vec);
            if (d.field == TF_FIELD_J) w = (1.0 / norm(ev)) * w;
            tf_source_term t{};
            t.edge = (int32_t)e;
            t.source = src_index[sl];
            t.weight = w;
            b.terms.push_back(t);
        }
        b.src_edge_desc.push_back(d);
    };
    for (int32_t e : b.uncommon_edges) add_edge((uint32_t)e);
    for (int32_t e : b.src_edges) add_edge((uint32_t)e);

    b.prob.method = TF_METHOD_FDTD;
    b.prob.n_faces = (int32_t)totF;
    b.prob.n_vols = (int32_t)nv;
    b.prob.duration = sim.duration;
    b.n_pts = (int32_t)np;
    b.first_step = 0;
    // for (i=0; i*t_step <= simulation_time; ++i)  (D:1840)
    int64_t i = 0;
    while ((double)i * t_step <= sim.duration) ++i;
    b.n_steps = i;
    clk.mark("sources");
}

}  // namespace tfh
