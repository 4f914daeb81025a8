// Mesh.cxx -- connectivity I/O and device staging (src/Mesh.cxx:11-103).
#include "Mesh.hxx"

#include <algorithm>
#include <cstdio>

Mesh::~Mesh() {
    if (!gpu_) return;
    for (int *p : {o_epoint, o_ef, o_fc, o_lines, o_linesize, o_lineface, o_linepoint}) gpu_->free(p);
}

void Mesh::finishRead() {
    nbytes = epoint.size() + ef.size() + fc.size() + linesize.size() + linepoint.size() +
             lines.size() + lineface.size();
    max_line_nelem = 0;
    for (int v : linesize) max_line_nelem = std::max(max_line_nelem, v);
}

bool Mesh::fromFile(const std::string &path) {
    int h[6];
    if (tbs_read_mesh_header(path.c_str(), h) != TBS_OK) {
        fprintf(stderr, "ERROR: %s\n", tbs_last_error());
        return false;
    }
    nvar = h[0]; nelem = h[1]; nintface = h[2]; eftot = h[3]; nline = h[4]; nlineelem = h[5];
    epoint.resize(nelem + 1); ef.resize(eftot); fc.resize(2 * (size_t)nintface);
    linesize.resize(nline); linepoint.resize(nline + 1); lines.resize(nlineelem);
    lineface.resize(nlineelem);
    if (tbs_read_mesh(path.c_str(), epoint.data(), ef.data(), fc.data(), linesize.data(),
                      linepoint.data(), lines.data(), lineface.data()) != TBS_OK) {
        fprintf(stderr, "ERROR: %s\n", tbs_last_error());
        return false;
    }
    // every index the kernels follow, before any of them can run
    const tbs_mesh_host mh = host();
    if (tbs_mesh_validate(&mh) != TBS_OK) {
        fprintf(stderr, "ERROR: %s: %s\n", path.c_str(), tbs_last_error());
        return false;
    }
    finishRead();
    printf("  Mesh Statistics:\n    nvar: %d\n    nelem: %d\n    nintface: %d\n    eftot: %d\n"
           "    nline: %d\n    nlineelem: %d\n  Max Line Element count: %d\n",
           nvar, nelem, nintface, eftot, nline, nlineelem, max_line_nelem);
    return true;
}

bool Mesh::fromStructured(int NL, int NK, int nv) {
    if (tbs_structured_sizes(NL, NK, &nelem, &nintface, &eftot, &nline, &nlineelem) != TBS_OK) return false;
    nvar = nv;
    epoint.resize(nelem + 1); ef.resize(eftot); fc.resize(2 * (size_t)nintface);
    linesize.resize(nline); linepoint.resize(nline + 1); lines.resize(nlineelem);
    lineface.resize(nlineelem);
    if (tbs_structured_mesh(NL, NK, epoint.data(), ef.data(), fc.data(), linesize.data(),
                            linepoint.data(), lines.data(), lineface.data()) != TBS_OK)
        return false;
    finishRead();
    return true;
}

bool Mesh::toFile(const std::string &path) const {
    int h[6] = {nvar, nelem, nintface, eftot, nline, nlineelem};
    return tbs_write_mesh(path.c_str(), h, epoint.data(), ef.data(), fc.data(), linesize.data(),
                          linepoint.data(), lines.data(), lineface.data()) == TBS_OK;
}

void Mesh::setupDevice(Platform &gpu) {
    gpu_ = &gpu;
    o_epoint = gpu.malloc<int>(epoint.size());
    o_ef = gpu.malloc<int>(ef.size());
    o_fc = gpu.malloc<int>(fc.size());
    o_lines = gpu.malloc<int>(lines.size());
    o_linesize = gpu.malloc<int>(linesize.size());
    o_lineface = gpu.malloc<int>(lineface.size());
    o_linepoint = gpu.malloc<int>(linepoint.size());
}

void Mesh::toDevice() {
    auto up = [&](int *d, const std::vector<int> &h) {
        tbs_check(tbs_memcpy_h2d(gpu_->ctx, d, h.data(), h.size() * sizeof(int)), "Mesh::toDevice");
    };
    up(o_epoint, epoint); up(o_ef, ef); up(o_fc, fc); up(o_lines, lines);
    up(o_linesize, linesize); up(o_lineface, lineface); up(o_linepoint, linepoint);
}

void Mesh::fromDevice() {
    auto down = [&](std::vector<int> &h, const int *d) {
        tbs_check(tbs_memcpy_d2h(gpu_->ctx, h.data(), d, h.size() * sizeof(int)), "Mesh::fromDevice");
    };
    down(epoint, o_epoint); down(ef, o_ef); down(fc, o_fc); down(lines, o_lines);
    down(linesize, o_linesize); down(lineface, o_lineface); down(linepoint, o_linepoint);
}

tbs_mesh_host Mesh::host() const {
    tbs_mesh_host m{};
    m.nvar = nvar; m.nelem = nelem; m.nintface = nintface; m.eftot = eftot;
    m.nline = nline; m.nlineelem = nlineelem;
    m.epoint = epoint.data(); m.ef = ef.data(); m.fc = fc.data(); m.linesize = linesize.data();
    m.linepoint = linepoint.data(); m.lines = lines.data(); m.lineface = lineface.data();
    return m;
}
