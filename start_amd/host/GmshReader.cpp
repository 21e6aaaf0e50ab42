// Gmsh 4.1 reader (ASCII and binary): nodes, 4-node tetrahedra and 3-node
// triangles, as the reference reads them through libgmsh
// (VolumetricMesh.cpp:35-121, GmshUtils.cpp:96-152).  Written from the Gmsh
// 4.1 file-format description: sections $MeshFormat, $Nodes and $Elements
// are read, every other section ($PhysicalNames, $Entities,
// $PartitionedEntities, ...) is skipped to its end marker.
#include "start_mesh.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_error;

// nodes of each linear and quadratic element type of the 4.1 format
int nodes_per_element(int type)
{
    switch (type) {
    case 1: return 2;    // line
    case 2: return 3;    // triangle
    case 3: return 4;    // quadrangle
    case 4: return 4;    // tetrahedron
    case 5: return 8;    // hexahedron
    case 6: return 6;    // prism
    case 7: return 5;    // pyramid
    case 8: return 3;    // line, 3 nodes
    case 9: return 6;    // triangle, 6 nodes
    case 10: return 9;   // quadrangle, 9 nodes
    case 11: return 10;  // tetrahedron, 10 nodes
    case 12: return 27;  // hexahedron, 27 nodes
    case 13: return 18;  // prism, 18 nodes
    case 14: return 14;  // pyramid, 14 nodes
    case 15: return 1;   // point
    case 16: return 8;   // quadrangle, 8 nodes
    case 17: return 20;  // hexahedron, 20 nodes
    case 18: return 15;  // prism, 15 nodes
    case 19: return 13;  // pyramid, 13 nodes
    default: return -1;
    }
}

struct Raw {
    std::vector<uint64_t> node_tag;
    std::vector<double> xyz;
    std::vector<uint64_t> tet_tag, tet_nodes, tri_tag, tri_nodes;
};

// --- binary helpers ---------------------------------------------------------
struct Bin {
    std::istream& in;
    template <typename T>
    T get()
    {
        T v;
        in.read(reinterpret_cast<char*>(&v), sizeof(T));
        if (!in) throw std::runtime_error("truncated binary section");
        return v;
    }
    void skip(std::streamoff n) { in.seekg(n, std::ios::cur); }
};

void skip_to(std::istream& in, const std::string& marker)
{
    // marker on a line of its own (also after binary data: scan by lines)
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line == marker) return;
    }
    throw std::runtime_error("missing " + marker);
}

void read_nodes_ascii(std::istream& in, Raw& r)
{
    size_t nblocks, nnodes, mn, mx;
    in >> nblocks >> nnodes >> mn >> mx;
    r.node_tag.reserve(nnodes);
    r.xyz.reserve(3 * nnodes);
    for (size_t b = 0; b < nblocks; ++b) {
        int dim, tag, parametric;
        size_t n;
        in >> dim >> tag >> parametric >> n;
        if (!in) throw std::runtime_error("bad $Nodes block header");
        std::vector<uint64_t> tags(n);
        for (size_t i = 0; i < n; ++i) in >> tags[i];
        const int extra = parametric ? dim : 0;
        for (size_t i = 0; i < n; ++i) {
            double x, y, z, u;
            in >> x >> y >> z;
            for (int e = 0; e < extra; ++e) in >> u;
            r.node_tag.push_back(tags[i]);
            r.xyz.push_back(x);
            r.xyz.push_back(y);
            r.xyz.push_back(z);
        }
        if (!in) throw std::runtime_error("bad $Nodes block");
    }
}

void read_nodes_binary(std::istream& in, Raw& r)
{
    Bin b{in};
    const uint64_t nblocks = b.get<uint64_t>(), nnodes = b.get<uint64_t>();
    (void)b.get<uint64_t>();
    (void)b.get<uint64_t>();
    r.node_tag.reserve(nnodes);
    r.xyz.reserve(3 * nnodes);
    for (uint64_t k = 0; k < nblocks; ++k) {
        const int dim = b.get<int32_t>();
        (void)b.get<int32_t>();
        const int parametric = b.get<int32_t>();
        const uint64_t n = b.get<uint64_t>();
        std::vector<uint64_t> tags(n);
        for (uint64_t i = 0; i < n; ++i) tags[i] = b.get<uint64_t>();
        const int extra = parametric ? dim : 0;
        for (uint64_t i = 0; i < n; ++i) {
            r.node_tag.push_back(tags[i]);
            for (int q = 0; q < 3; ++q) r.xyz.push_back(b.get<double>());
            b.skip((std::streamoff)(sizeof(double) * extra));
        }
    }
}

void keep_element(Raw& r, int type, uint64_t tag, const uint64_t* nodes)
{
    if (type == 4) {
        r.tet_tag.push_back(tag);
        r.tet_nodes.insert(r.tet_nodes.end(), nodes, nodes + 4);
    } else if (type == 2) {
        r.tri_tag.push_back(tag);
        r.tri_nodes.insert(r.tri_nodes.end(), nodes, nodes + 3);
    }
}

void read_elements_ascii(std::istream& in, Raw& r)
{
    size_t nblocks, nelem, mn, mx;
    in >> nblocks >> nelem >> mn >> mx;
    for (size_t b = 0; b < nblocks; ++b) {
        int dim, tag, type;
        size_t n;
        in >> dim >> tag >> type >> n;
        if (!in) throw std::runtime_error("bad $Elements block header");
        const int npe = nodes_per_element(type);
        if (npe < 0) throw std::runtime_error("unknown element type " + std::to_string(type));
        std::vector<uint64_t> nodes((size_t)npe);
        for (size_t i = 0; i < n; ++i) {
            uint64_t et;
            in >> et;
            for (int q = 0; q < npe; ++q) in >> nodes[(size_t)q];
            keep_element(r, type, et, nodes.data());
        }
        if (!in) throw std::runtime_error("bad $Elements block");
    }
}

void read_elements_binary(std::istream& in, Raw& r)
{
    Bin b{in};
    const uint64_t nblocks = b.get<uint64_t>();
    (void)b.get<uint64_t>();
    (void)b.get<uint64_t>();
    (void)b.get<uint64_t>();
    for (uint64_t k = 0; k < nblocks; ++k) {
        (void)b.get<int32_t>();
        (void)b.get<int32_t>();
        const int type = b.get<int32_t>();
        const uint64_t n = b.get<uint64_t>();
        const int npe = nodes_per_element(type);
        if (npe < 0) throw std::runtime_error("unknown element type " + std::to_string(type));
        std::vector<uint64_t> rec((size_t)npe + 1);
        for (uint64_t i = 0; i < n; ++i) {
            for (int q = 0; q <= npe; ++q) rec[(size_t)q] = b.get<uint64_t>();
            keep_element(r, type, rec[0], rec.data() + 1);
        }
    }
}

void read_file(const char* path, Raw& r)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error(std::string("cannot open ") + path);
    std::string line;
    bool binary = false, have_format = false, have_nodes = false, have_elements = false;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty()) continue;
        if (line == "$MeshFormat") {
            std::getline(in, line);
            std::istringstream ss(line);
            double version;
            int file_type, dsize;
            ss >> version >> file_type >> dsize;
            if (!ss || version < 4.1 || version >= 5.0)
                throw std::runtime_error("only the Gmsh 4.1 format is read (found '" + line + "')");
            if (dsize != 8) throw std::runtime_error("data size must be 8");
            binary = file_type == 1;
            if (binary) {
                int32_t one = 0;
                in.read(reinterpret_cast<char*>(&one), sizeof(one));
                if (one != 1) throw std::runtime_error("binary file of the other endianness");
            }
            skip_to(in, "$EndMeshFormat");
            have_format = true;
        } else if (line == "$Nodes") {
            if (!have_format) throw std::runtime_error("$Nodes before $MeshFormat");
            binary ? read_nodes_binary(in, r) : read_nodes_ascii(in, r);
            skip_to(in, "$EndNodes");
            have_nodes = true;
        } else if (line == "$Elements") {
            if (!have_format) throw std::runtime_error("$Elements before $MeshFormat");
            binary ? read_elements_binary(in, r) : read_elements_ascii(in, r);
            skip_to(in, "$EndElements");
            have_elements = true;
        } else if (line[0] == '$' && line.compare(0, 4, "$End") != 0) {
            skip_to(in, "$End" + line.substr(1));
        }
    }
    if (!have_nodes || !have_elements) throw std::runtime_error("no $Nodes or $Elements section");
}

template <typename T>
T* copy_out(const std::vector<T>& v)
{
    T* p = static_cast<T*>(std::malloc(sizeof(T) * std::max<size_t>(v.size(), 1)));
    if (!p) throw std::runtime_error("out of host memory");
    if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

}  // namespace

extern "C" {

const char* start_gmsh_last_error(void) { return g_error.c_str(); }

void start_gmsh_free(start_gmsh_mesh* m)
{
    if (!m) return;
    void* ptrs[] = {m->nodes, m->node_tags, m->tets, m->tet_tags, m->tris, m->tri_tags};
    for (void* p : ptrs) std::free(p);
    *m = start_gmsh_mesh{};
}

int start_gmsh_read(const char* path, start_gmsh_mesh* out)
{
    if (!path || !out) {
        g_error = "null argument";
        return -1;
    }
    *out = start_gmsh_mesh{};
    try {
        Raw r;
        read_file(path, r);
        const size_t nn = r.node_tag.size();
        // nodes in ascending tag order; node references become rows
        std::vector<size_t> order(nn);
        for (size_t i = 0; i < nn; ++i) order[i] = i;
        std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return r.node_tag[a] < r.node_tag[b]; });
        std::unordered_map<uint64_t, int32_t> row;
        row.reserve(nn * 2);
        std::vector<double> xyz(3 * nn);
        std::vector<uint64_t> tags(nn);
        for (size_t k = 0; k < nn; ++k) {
            const size_t i = order[k];
            if (k && r.node_tag[i] == tags[k - 1]) throw std::runtime_error("duplicate node tag " + std::to_string(tags[k - 1]));
            tags[k] = r.node_tag[i];
            row[tags[k]] = (int32_t)k;
            for (int q = 0; q < 3; ++q) xyz[3 * k + q] = r.xyz[3 * i + q];
        }
        auto rows = [&](const std::vector<uint64_t>& refs) {
            std::vector<int32_t> v(refs.size());
            for (size_t i = 0; i < refs.size(); ++i) {
                auto it = row.find(refs[i]);
                if (it == row.end()) throw std::runtime_error("element refers to unknown node tag " + std::to_string(refs[i]));
                v[i] = it->second;
            }
            return v;
        };
        std::vector<int32_t> tets = rows(r.tet_nodes), tris = rows(r.tri_nodes);
        // Elements in ascending tag order: the device breaks owner ties (a point
        // on a shared face) by the smaller row and hit ties by the smaller
        // triangle row, where the reference takes the smaller element tag
        // (its maps are keyed by tag: NodeChargeDensityProcessor.cpp:51-56).
        auto by_tag = [](std::vector<uint64_t>& tag, std::vector<int32_t>& conn, int per, const char* what) {
            std::vector<size_t> ord(tag.size());
            for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
            std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return tag[a] < tag[b]; });
            std::vector<uint64_t> t2(tag.size());
            std::vector<int32_t> c2(conn.size());
            for (size_t k = 0; k < ord.size(); ++k) {
                t2[k] = tag[ord[k]];
                if (k && t2[k] == t2[k - 1])
                    throw std::runtime_error(std::string("duplicate ") + what + " tag " + std::to_string(t2[k]));
                for (int q = 0; q < per; ++q) c2[per * k + q] = conn[per * ord[k] + q];
            }
            tag.swap(t2);
            conn.swap(c2);
        };
        by_tag(r.tet_tag, tets, 4, "tetrahedron");
        by_tag(r.tri_tag, tris, 3, "triangle");
        out->n_nodes = (int64_t)nn;
        out->nodes = copy_out(xyz);
        out->node_tags = copy_out(tags);
        out->n_tets = (int64_t)r.tet_tag.size();
        out->tets = copy_out(tets);
        out->tet_tags = copy_out(r.tet_tag);
        out->n_tris = (int64_t)r.tri_tag.size();
        out->tris = copy_out(tris);
        out->tri_tags = copy_out(r.tri_tag);
    } catch (const std::exception& e) {
        start_gmsh_free(out);
        g_error = std::string(path) + ": " + e.what();
        return -1;
    }
    return 0;
}

}  // extern "C"
