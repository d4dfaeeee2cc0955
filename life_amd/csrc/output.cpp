// Result files of a life_case in the reference's formats: restart
// (Fluid/IBM/FEM.restart), VTK (Fluid.<t>.vti, IBM.<t>.vtp, FEM.<t>.vtp) and
// the force / tip logs, under a results directory laid out as the
// reference's Results/ folder (Restart/, VTK/, *.out).
//
// Binary data is written in the host's byte order, as the reference does on
// a little-endian host (its swapEnd applies on big-endian hosts only,
// Utils.h:109-135); text uses iostreams with the reference's precisions, so
// the bytes are the same as the reference's for the same state.
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "case_impl.h"

using namespace life;
namespace fs = std::filesystem;

namespace {

constexpr int kPrecision = 10;  // PRECISION, params.h:110

template <typename F>
int guarded_io(F&& f) {
    try {
        f();
        return 0;
    } catch (const HipError& e) {
        set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
        return 2;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    }
}

template <typename T>
void put(std::ofstream& o, T v) {
    o.write(reinterpret_cast<const char*>(&v), sizeof(T));
}

template <typename T>
T get(std::ifstream& in, const std::string& what) {
    T v;
    in.read(reinterpret_cast<char*>(&v), sizeof(T));
    if (!in) throw std::runtime_error("truncated " + what);
    return v;
}

std::ofstream open_out(const std::string& name, const std::string& what) {
    std::ofstream o(name, std::ios::binary);
    if (!o.is_open()) throw std::runtime_error("Error opening " + what + " file...exiting");
    return o;
}

const char* endian_str() {
    const uint32_t probe = 0x01020304u;
    return reinterpret_cast<const unsigned char*>(&probe)[0] == 1 ? "BigEndian" : "LittleEndian";
}

// Lattice state in the reference layout (the arrays GridClass writes).
struct GridState {
    std::vector<double> f, rho, u, fibm;
};

GridState download_grid(life_case* c) {
    const size_t N = (size_t)c->p.Nx * c->p.Ny;
    GridState s;
    s.f.resize(9 * N);
    s.rho.resize(N);
    s.u.resize(2 * N);
    s.fibm.resize(2 * N);
    if (life_grid_download(c->g, s.f.data(), s.rho.data(), s.u.data(), s.fibm.data()))
        throw std::runtime_error(life_last_error());
    return s;
}

// Marker positions, velocities and forces on the host (device copies are
// the last step's final values).
void download_markers(life_case* c) {
    life_grid* g = c->g;
    if (!c->nm) return;
    LIFE_HIP(hipSetDevice(g->device));
    LIFE_HIP(hipEventSynchronize(c->ev_upload));
    LIFE_HIP(hipMemcpyAsync(c->pos.data(), g->md.pos, sizeof(double) * 2 * c->nm, hipMemcpyDeviceToHost, g->stream));
    LIFE_HIP(hipMemcpyAsync(c->vel.data(), g->md.vel, sizeof(double) * 2 * c->nm, hipMemcpyDeviceToHost, g->stream));
    LIFE_HIP(hipMemcpyAsync(c->force.data(), g->md.force, sizeof(double) * 2 * c->nm, hipMemcpyDeviceToHost,
                            g->stream));
    LIFE_HIP(hipStreamSynchronize(g->stream));
}

// ---------------------------------------------------------------- restart

// GridClass::writeRestart, Grid.cpp:1163-1229
void write_fluid_restart(life_case* c, const std::string& dir) {
    const GridState s = download_grid(c);
    const CaseScalars& cs = c->cs;
    const std::string tmp = dir + "/Fluid.restart.temp";
    {
        std::ofstream o = open_out(tmp, "Fluid.restart.temp");
        put<int>(o, c->t);
        put<int>(o, c->p.Nx);
        put<int>(o, c->p.Ny);
        put<double>(o, cs.omega);
        put<double>(o, cs.Dx);
        put<double>(o, cs.Dt);
        put<double>(o, cs.Dm);
        const int Ny = c->p.Ny;
        for (int i = 0; i < c->p.Nx; i++)
            for (int j = 0; j < Ny; j++) {
                const size_t id = (size_t)i * Ny + j;
                put<int>(o, i);
                put<int>(o, j);
                put<double>(o, s.rho[id]);
                put<double>(o, s.u[2 * id]);
                put<double>(o, s.u[2 * id + 1]);
                put<double>(o, s.fibm[2 * id]);
                put<double>(o, s.fibm[2 * id + 1]);
                o.write(reinterpret_cast<const char*>(&s.f[9 * id]), 9 * sizeof(double));
            }
    }
    fs::rename(tmp, dir + "/Fluid.restart");
}

// GridClass::readRestart, Grid.cpp:1072-1160
void read_fluid_restart(life_case* c, const std::string& dir) {
    std::ifstream in(dir + "/Fluid.restart", std::ios::binary);
    if (!in.is_open()) throw std::runtime_error("Error opening Fluid.restart file...exiting");
    const int t = get<int>(in, "Fluid.restart");
    const int nx = get<int>(in, "Fluid.restart"), ny = get<int>(in, "Fluid.restart");
    const double omega = get<double>(in, "Fluid.restart"), dx = get<double>(in, "Fluid.restart");
    const double dt = get<double>(in, "Fluid.restart"), dm = get<double>(in, "Fluid.restart");
    const CaseScalars& cs = c->cs;
    if (c->p.Nx != nx || c->p.Ny != ny || cs.omega != omega || cs.Dx != dx || cs.Dt != dt || cs.Dm != dm)
        throw std::runtime_error("Grid size/scaling has changed between runs...this is not supported");
    const size_t N = (size_t)nx * ny;
    GridState s;
    s.f.resize(9 * N);
    s.rho.resize(N);
    s.u.resize(2 * N);
    s.fibm.resize(2 * N);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            const size_t id = (size_t)i * ny + j;
            const int ir = get<int>(in, "Fluid.restart"), jr = get<int>(in, "Fluid.restart");
            if (ir != i || jr != j) throw std::runtime_error("Grid indices do not match Fluid.restart file...exiting");
            s.rho[id] = get<double>(in, "Fluid.restart");
            s.u[2 * id] = get<double>(in, "Fluid.restart");
            s.u[2 * id + 1] = get<double>(in, "Fluid.restart");
            s.fibm[2 * id] = get<double>(in, "Fluid.restart");
            s.fibm[2 * id + 1] = get<double>(in, "Fluid.restart");
            in.read(reinterpret_cast<char*>(&s.f[9 * id]), 9 * sizeof(double));
            if (!in) throw std::runtime_error("truncated Fluid.restart");
        }
    c->g->g = step_scalars(*c->g, t);  // force_xy of the restart step (WOMERSLEY)
    c->g->t_step = t;
    if (life_grid_upload(c->g, s.f.data(), s.rho.data(), s.u.data(), s.fibm.data()))
        throw std::runtime_error(life_last_error());
    c->t = t;
}

// ObjectsClass::writeRestart, Objects.cpp:1206-1386
void write_objects_restart(life_case* c, const std::string& dir) {
    if (c->hasIBM) {
        download_markers(c);
        const std::string tmp = dir + "/IBM.restart.temp";
        {
            std::ofstream o = open_out(tmp, "IBM.restart.temp");
            put<size_t>(o, c->bodies.size());
            for (const Body& b : c->bodies) {
                put<int>(o, b.ID);
                put<size_t>(o, (size_t)b.n);
                for (int n = 0; n < b.n; n++) {
                    const int k = b.first + n;
                    put<double>(o, c->pos[2 * k]);
                    put<double>(o, c->pos[2 * k + 1]);
                    put<double>(o, c->vel[2 * k]);
                    put<double>(o, c->vel[2 * k + 1]);
                    put<double>(o, c->force[2 * k]);
                    put<double>(o, c->force[2 * k + 1]);
                }
            }
        }
        fs::rename(tmp, dir + "/IBM.restart");
    }
    if (c->hasFlex) {
        const std::string tmp = dir + "/FEM.restart.temp";
        {
            std::ofstream o = open_out(tmp, "FEM.restart.temp");
            put<int>(o, (int)c->flex_bodies.size());
            put<double>(o, c->relax);
            for (const Body* b : c->flex_bodies) {
                const FemBody& f = *b->fem;
                put<int>(o, b->ID);
                put<size_t>(o, (size_t)f.num_elements());
                put<size_t>(o, (size_t)f.num_nodes());
                for (const std::vector<double>* v : {&f.U, &f.Ud, &f.Udd, &f.Un, &f.Unm1, &f.Rk})
                    o.write(reinterpret_cast<const char*>(v->data()), sizeof(double) * f.ndof());
                for (int e = 0; e < f.num_elements(); e++) {
                    double L, a;
                    f.element_state(e, L, a);
                    put<double>(o, L);
                    put<double>(o, a);
                }
                for (int n = 0; n < f.num_nodes(); n++) {
                    double v[6];
                    f.node_state(n, v);
                    o.write(reinterpret_cast<const char*>(v), sizeof v);
                }
            }
        }
        fs::rename(tmp, dir + "/FEM.restart");
    }
}

int body_index(const life_case* c, int id) {
    for (size_t b = 0; b < c->bodies.size(); b++)
        if (c->bodies[b].ID == id) return (int)b;
    return -1;
}

// ObjectsClass::readRestart, Objects.cpp:979-1204 (file part): bodies
// missing from this run are skipped, node counts must match.
void read_objects_restart(life_case* c, const std::string& dir) {
    const std::string ibm = dir + "/IBM.restart", fem = dir + "/FEM.restart";
    if (fs::exists(ibm)) {
        std::ifstream in(ibm, std::ios::binary);
        if (!in.is_open()) throw std::runtime_error("Error opening IBM.restart file...exiting");
        const size_t nb = get<size_t>(in, "IBM.restart");
        for (size_t q = 0; q < nb; q++) {
            const int id = get<int>(in, "IBM.restart");
            const size_t nn = get<size_t>(in, "IBM.restart");
            const int b = body_index(c, id);
            if (b < 0) {
                in.seekg((std::streamoff)(6 * sizeof(double) * nn), std::ios::cur);
                continue;
            }
            const Body& body = c->bodies[b];
            if (nn != (size_t)body.n)
                throw std::runtime_error(
                    "Number of nodes in IBM.restart do not match existing number of nodes...exiting");
            for (int n = 0; n < body.n; n++) {
                const int k = body.first + n;
                double v[6];
                in.read(reinterpret_cast<char*>(v), sizeof v);
                if (!in) throw std::runtime_error("truncated IBM.restart");
                c->pos[2 * k] = v[0];
                c->pos[2 * k + 1] = v[1];
                c->vel[2 * k] = v[2];
                c->vel[2 * k + 1] = v[3];
                c->force[2 * k] = v[4];
                c->force[2 * k + 1] = v[5];
            }
        }
    }
    if (fs::exists(fem)) {
        std::ifstream in(fem, std::ios::binary);
        if (!in.is_open()) throw std::runtime_error("Error opening FEM.restart file...exiting");
        const int nf = get<int>(in, "FEM.restart");
        c->relax = get<double>(in, "FEM.restart");
        for (int q = 0; q < nf; q++) {
            const int id = get<int>(in, "FEM.restart");
            const size_t nel = get<size_t>(in, "FEM.restart"), nnode = get<size_t>(in, "FEM.restart");
            const int b = body_index(c, id);
            if (b < 0 || c->bodies[b].flex != 0) {
                const size_t dofs = 3 * nnode;  // nodeDOFs
                in.seekg((std::streamoff)(6 * sizeof(double) * dofs + 2 * sizeof(double) * nel +
                                          6 * sizeof(double) * nnode),
                         std::ios::cur);
                continue;
            }
            FemBody& f = *c->bodies[b].fem;
            if (nel != (size_t)f.num_elements() || nnode != (size_t)f.num_nodes())
                throw std::runtime_error("Number of nodes/elements in FEM.restart do not match existing number of "
                                         "nodes/elements...exiting");
            for (std::vector<double>* v : {&f.U, &f.Ud, &f.Udd, &f.Un, &f.Unm1, &f.Rk}) {
                in.read(reinterpret_cast<char*>(v->data()), sizeof(double) * f.ndof());
                if (!in) throw std::runtime_error("truncated FEM.restart");
            }
            for (int e = 0; e < f.num_elements(); e++) {
                const double L = get<double>(in, "FEM.restart"), a = get<double>(in, "FEM.restart");
                f.set_element_state(e, L, a);
            }
            for (int n = 0; n < f.num_nodes(); n++) {
                double v[6];
                in.read(reinterpret_cast<char*>(v), sizeof v);
                if (!in) throw std::runtime_error("truncated FEM.restart");
                f.set_node_state(n, v);
            }
        }
    }
}

// ---------------------------------------------------------------- VTK

// ObjectsClass::writeEmptyVTK, Objects.cpp:1462-1509
void write_empty_vtp(const std::string& name) {
    std::ofstream o = open_out(name, "body VTK");
    o << "<?xml version=\"1.0\"?>\n";
    o << "<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"" << endian_str() << "\" header_type=\"UInt64\">\n";
    o << "\t<PolyData>\n";
    o << "\t\t<Piece NumberOfPoints=\"0\" NumberOfVerts=\"0\" NumberOfLines=\"0\" NumberOfStrips=\"0\" "
         "NumberOfPolys=\"0\">\n";
    o << "\t\t</Piece>\n";
    o << "\t</PolyData>\n";
    o << "</VTKFile>\n";
}

// GridClass::writeVTK, Grid.cpp:790-913: density, pressure and velocity in
// physical units on the image grid, j-major.
void write_fluid_vtk(life_case* c, const std::string& vdir) {
    const GridState s = download_grid(c);
    const CaseScalars& cs = c->cs;
    const life_params& p = c->p;
    const int Nx = p.Nx, Ny = p.Ny;
    std::ofstream o = open_out(vdir + "/Fluid." + std::to_string(c->t) + ".vti", "fluid VTK");
    const unsigned long long bytes = (unsigned long long)Nx * Ny * sizeof(double);
    o << "<?xml version=\"1.0\"?>\n";
    o << "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"" << endian_str() << "\" header_type=\"UInt64\">\n";
    o << "\t<ImageData WholeExtent=\"" << 0.0 << " " << Nx - 1 << " " << 0.0 << " " << Ny - 1 << " " << 0.0 << " "
      << 0.0 << "\" Origin=\"" << 0.0 << " " << 0.0 << " " << 0.0 << "\" Spacing=\"" << cs.Dx << " " << cs.Dx << " "
      << cs.Dx << "\">\n";
    o << "\t\t<Piece Extent=\"" << 0.0 << " " << Nx - 1 << " " << 0.0 << " " << Ny - 1 << " " << 0.0 << " " << 0.0
      << "\">\n";
    o << "\t\t\t<PointData>\n";
    o << "\t\t\t\t<DataArray type=\"Float64\" Name=\"Density\" format=\"appended\" offset=\"" << 0 << "\"/>\n";
    o << "\t\t\t\t<DataArray type=\"Float64\" Name=\"Pressure\" format=\"appended\" offset=\""
      << 1 * (bytes + sizeof(unsigned long long)) << "\"/>\n";
    o << "\t\t\t\t<DataArray type=\"Float64\" Name=\"Velocity\" NumberOfComponents=\"3\" format=\"appended\" offset=\""
      << 2 * (bytes + sizeof(unsigned long long)) << "\"/>\n";
    o << "\t\t\t</PointData>\n";
    o << "\t\t</Piece>\n";
    o << "\t</ImageData>\n";
    o << "\t<AppendedData encoding=\"raw\">\n";
    o << "\t\t_";
    std::vector<double> buf((size_t)Nx * 3);
    put<unsigned long long>(o, bytes);
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) buf[i] = s.rho[(size_t)i * Ny + j] * cs.Drho;
        o.write(reinterpret_cast<const char*>(buf.data()), sizeof(double) * Nx);
    }
    put<unsigned long long>(o, bytes);
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) buf[i] = p.ref_P + (s.rho[(size_t)i * Ny + j] - p.rho_p / cs.Drho) * cs.cs2 * cs.Dm /
                                                            (cs.Dx * (cs.Dt * cs.Dt));
        o.write(reinterpret_cast<const char*>(buf.data()), sizeof(double) * Nx);
    }
    put<unsigned long long>(o, 3 * bytes);
    const double vscale = cs.Dx / cs.Dt;
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) {
            const size_t id = (size_t)i * Ny + j;
            buf[3 * i] = s.u[2 * id] * vscale;
            buf[3 * i + 1] = s.u[2 * id + 1] * vscale;
            buf[3 * i + 2] = 0.0;
        }
        o.write(reinterpret_cast<const char*>(buf.data()), sizeof(double) * 3 * Nx);
    }
    o << "\n\t</AppendedData>\n";
    o << "</VTKFile>\n";
}

// ObjectsClass::writeVTK, Objects.cpp:711-864: one piece per body (IBM:
// markers; FEM: beam nodes of flexible bodies), closed polyline for
// circles.
void write_body_vtk(life_case* c, const std::string& vdir, bool ibm) {
    std::vector<const Body*> bodies;
    for (const Body& b : c->bodies)
        if (ibm || b.flex == 0) bodies.push_back(&b);
    std::vector<std::vector<double>> pts(bodies.size());
    for (size_t q = 0; q < bodies.size(); q++) {
        const Body& b = *bodies[q];
        if (ibm) {
            for (int n = 0; n < b.n; n++) {
                pts[q].push_back(c->pos[2 * (b.first + n)]);
                pts[q].push_back(c->pos[2 * (b.first + n) + 1]);
            }
        } else {
            for (int n = 0; n < b.fem->num_nodes(); n++) {
                double v[6];
                b.fem->node_state(n, v);
                pts[q].push_back(v[2]);
                pts[q].push_back(v[3]);
            }
        }
    }
    const std::string name = std::string(ibm ? "IBM" : "FEM") + "." + std::to_string(c->t) + ".vtp";
    std::ofstream o = open_out(vdir + "/" + name, "body VTK");
    o << "<?xml version=\"1.0\"?>\n";
    o << "<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"" << endian_str() << "\" header_type=\"UInt64\">\n";
    o << "\t<PolyData>\n";
    unsigned long long offset = 0;
    for (size_t q = 0; q < bodies.size(); q++) {
        const size_t nn = pts[q].size() / 2;
        const size_t nl = bodies[q]->kind == 0 ? nn : nn - 1;
        o << "\t\t<Piece NumberOfPoints=\"" << nn << "\" NumberOfVerts=\"" << 0 << "\" NumberOfLines=\"" << nl
          << "\" NumberOfStrips=\"" << 0 << "\" NumberOfPolys=\"" << 0 << "\">\n";
        o << "\t\t\t<Points>\n";
        o << "\t\t\t\t<DataArray type=\"Float64\" Name=\"Points\" NumberOfComponents=\"3\" format=\"appended\" offset=\""
          << offset << "\"/>\n";
        offset += 3 * nn * sizeof(double) + sizeof(unsigned long long);
        o << "\t\t\t</Points>\n";
        o << "\t\t\t<Lines>\n";
        o << "\t\t\t\t<DataArray type=\"Int64\" Name=\"connectivity\" format=\"appended\" offset=\"" << offset
          << "\"/>\n";
        offset += 2 * nl * sizeof(long long) + sizeof(unsigned long long);
        o << "\t\t\t\t<DataArray type=\"Int64\" Name=\"offsets\" format=\"appended\" offset=\"" << offset << "\"/>\n";
        offset += nl * sizeof(long long) + sizeof(unsigned long long);
        o << "\t\t\t</Lines>\n";
        o << "\t\t</Piece>\n";
    }
    o << "\t</PolyData>\n";
    o << "\t<AppendedData encoding=\"raw\">\n";
    o << "\t\t_";
    for (size_t q = 0; q < bodies.size(); q++) {
        const size_t nn = pts[q].size() / 2;
        const size_t nl = bodies[q]->kind == 0 ? nn : nn - 1;
        put<unsigned long long>(o, 3 * nn * sizeof(double));
        for (size_t n = 0; n < nn; n++) {
            put<double>(o, pts[q][2 * n]);
            put<double>(o, pts[q][2 * n + 1]);
            put<double>(o, 0.0);
        }
        put<unsigned long long>(o, 2 * nl * sizeof(long long));
        for (size_t n = 0; n < nl; n++) {
            put<long long>(o, (long long)n);
            put<long long>(o, (long long)((n + 1) % nn));
        }
        put<unsigned long long>(o, nl * sizeof(long long));
        for (size_t n = 0; n < nl; n++) put<long long>(o, (long long)(2 * (n + 1)));
    }
    o << "\n\t</AppendedData>\n";
    o << "</VTKFile>\n";
}

// ---------------------------------------------------------------- logs

// Append one line to a tab-separated log; the header goes first when the
// file is new (Objects.cpp:497-513).
void append_line(const std::string& name, const char* header, const std::string& line) {
    const bool existing = fs::exists(name);
    std::ofstream o(name, std::ios::app);
    if (!o.is_open()) throw std::runtime_error("Error opening " + name + "...exiting");
    if (!existing) o << header;
    o << "\n" << line;
}

// ObjectsClass::deleteFileOutput, Objects.cpp:1525-1586: keep the header
// and the lines of timesteps before t.
void truncate_log(const std::string& name, int t, const std::string& dir) {
    if (!fs::exists(name)) return;
    std::ifstream in(name);
    if (!in.is_open()) throw std::runtime_error("Error opening file for deleting future tips/forces...exiting");
    std::vector<std::string> keep;
    std::string line;
    std::getline(in, line);
    keep.push_back(line);
    while (std::getline(in, line)) {
        if (std::stoi(line) >= t) break;
        keep.push_back(line);
    }
    in.close();
    const std::string tmp = dir + "/Output.temp";
    {
        std::ofstream o(tmp);
        if (!o.is_open()) throw std::runtime_error("Error opening temporary output file...exiting");
        o << keep[0];
        for (size_t k = 1; k < keep.size(); k++) o << "\n" << keep[k];
    }
    fs::rename(tmp, name);
}

std::string fmt_line(const life_case* c, const std::vector<double>& vals) {
    std::ostringstream s;
    s.precision(kPrecision);
    s << c->t << "\t" << c->cs.Dt * c->t;
    for (double v : vals) s << "\t" << v;
    return s.str();
}

}  // namespace

extern "C" {

int life_case_write_restart(life_case* c, const char* results) {
    return guarded_io([&] {
        case_sync_host(c);
        const std::string dir = std::string(results) + "/Restart";
        fs::create_directories(dir);
        write_fluid_restart(c, dir);
        write_objects_restart(c, dir);
    });
}

int life_case_read_restart(life_case* c, const char* results) {
    return guarded_io([&] {
        const std::string base(results), dir = base + "/Restart";
        read_fluid_restart(c, dir);
        if (c->hasIBM) {
            case_sync_host(c);  // nothing the device holds is newer than what the file replaces
            read_objects_restart(c, dir);
            case_restart_objects(c);
            case_push_fem(c);
        }
        // ObjectsClass::restartCleanup, Objects.cpp:1387-1405
        const std::string vdir = base + "/VTK";
        if (c->hasIBM && fs::exists(vdir)) {
            for (const auto& e : fs::directory_iterator(vdir)) {
                const std::string stem = e.path().stem().string();
                if (e.path().extension() == ".vti" && stem.rfind("Fluid.", 0) == 0) {
                    const std::string step = stem.substr(6);
                    const std::string body = vdir + "/IBM." + step + ".vtp";
                    if (!fs::exists(body)) write_empty_vtp(body);
                }
            }
        }
        for (const char* f : {"TipPositions.out", "TipVelocities.out", "TotalForces.out"})
            truncate_log(base + "/" + f, c->t, base);
        if (!c->hasIBM && fs::exists(dir + "/IBM.restart")) fs::remove(dir + "/IBM.restart");
        if (!c->hasFlex && fs::exists(dir + "/FEM.restart")) fs::remove(dir + "/FEM.restart");
    });
}

int life_case_write_vtk(life_case* c, const char* results, int fem) {
    return guarded_io([&] {
        const std::string vdir = std::string(results) + "/VTK";
        fs::create_directories(vdir);
        write_fluid_vtk(c, vdir);
        if (!c->hasIBM) {
            // Grid.cpp:903-912: bodies written earlier in the run get empty files
            if (fs::exists(vdir + "/IBM.0.vtp")) write_empty_vtp(vdir + "/IBM." + std::to_string(c->t) + ".vtp");
            if (fem && fs::exists(vdir + "/FEM.0.vtp"))
                write_empty_vtp(vdir + "/FEM." + std::to_string(c->t) + ".vtp");
            return;
        }
        download_markers(c);
        write_body_vtk(c, vdir, true);
        if (fem && c->hasFlex) write_body_vtk(c, vdir, false);
        else if (fem && fs::exists(vdir + "/FEM.0.vtp"))
            write_empty_vtp(vdir + "/FEM." + std::to_string(c->t) + ".vtp");
    });
}

int life_case_write_forces(life_case* c, const char* results) {
    return guarded_io([&] {
        if (!c->hasIBM) return;
        double F[4];
        if (life_case_total_force(c, F)) throw std::runtime_error(life_last_error());
        fs::create_directories(results);
        append_line(std::string(results) + "/TotalForces.out", "Timestep\tTime (s)\tFx (N)\tFy (N)\tCx\tCy",
                    fmt_line(c, {F[0], F[1], F[2], F[3]}));
    });
}

int life_case_write_tips(life_case* c, const char* results) {
    return guarded_io([&] {
        if (!c->hasFlex) return;
        download_markers(c);
        std::vector<double> pos, vel;
        for (const Body* b : c->flex_bodies) {
            const int k = b->first + b->n - 1;
            pos.push_back(c->pos[2 * k]);
            pos.push_back(c->pos[2 * k + 1]);
            vel.push_back(c->vel[2 * k]);
            vel.push_back(c->vel[2 * k + 1]);
        }
        fs::create_directories(results);
        append_line(std::string(results) + "/TipPositions.out", "Timestep\tTime (s)\tTipX\tTipY", fmt_line(c, pos));
        append_line(std::string(results) + "/TipVelocities.out", "Timestep\tTime (s)\tTipX\tTipY", fmt_line(c, vel));
    });
}

}  // extern "C"
