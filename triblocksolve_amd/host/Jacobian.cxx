// Jacobian.cxx -- Jacobian values, A packing and device staging
// (src/Jacobian.cxx:11-290).
#include "Jacobian.hxx"

#include <cstdio>
#include <algorithm>
#include <cstring>

Jacobian::~Jacobian() {
    if (!gpu_) return;
    for (double *p : {o_jacDinvC, o_jacDLU, o_jacD, o_jacO1, o_jacO2, o_rhs, o_res, o_U, o_U0, o_dU, o_A})
        gpu_->free(p);
}

void Jacobian::allocate(int nlineelem) {
    const size_t bs = (size_t)nvar * nvar;
    jacDLU.assign(bs * nelem, 0.0); jacD.assign(bs * nelem, 0.0);
    jacO1.assign(bs * nintface, 0.0); jacO2.assign(bs * nintface, 0.0);
    rhs.assign((size_t)nvar * nelem, 0.0); U0.assign((size_t)nvar * nelem, 0.0);
    U.assign((size_t)nvar * nelem, 0.0); dU.assign((size_t)nvar * nelem, 0.0);
    res.assign((size_t)nvar * nelem, 0.0);
    DinvC.assign(bs * nelem, 0.0);
    // the reference sizes A by nlineelem but indexes it by element id
    // (src/Jacobian.cxx:49,103); size it for both
    A.assign(bs * (size_t)std::max(nelem, nlineelem), 0.0);
    nbytes = jacD.size() + jacO1.size() + jacO2.size() + rhs.size() + U0.size() + U.size() +
             dU.size() + res.size();
}

bool Jacobian::fromFile(int nlineelem, const std::string &path) {
    int h[3];
    if (tbs_read_jacobian_header(path.c_str(), h) != TBS_OK) {
        fprintf(stderr, "ERROR: %s\n", tbs_last_error());
        return false;
    }
    nvar = h[0]; nelem = h[1]; nintface = h[2];
    printf("Read Jac Stats:\n  nvar: %d\n  nelem: %d\n  nintface: %d\n", nvar, nelem, nintface);
    allocate(nlineelem);
    if (tbs_read_jacobian(path.c_str(), jacD.data(), jacO1.data(), jacO2.data(), rhs.data(),
                          U0.data()) != TBS_OK) {
        fprintf(stderr, "ERROR: %s\n", tbs_last_error());
        return false;
    }
    return true;
}

bool Jacobian::fromGenerator(const Mesh &mesh, uint64_t seed, double shift) {
    nvar = mesh.nvar; nelem = mesh.nelem; nintface = mesh.nintface;
    allocate(mesh.nlineelem);
    return tbs_generate_host(seed, nvar, nelem, nintface, shift, jacD.data(), jacO1.data(),
                             jacO2.data(), rhs.data()) == TBS_OK;
}

bool Jacobian::toFile(const std::string &path) const {
    return tbs_write_jacobian(path.c_str(), nvar, nelem, nintface, jacD.data(), jacO1.data(),
                              jacO2.data(), rhs.data(), U0.data()) == TBS_OK;
}

void Jacobian::assembleTriBlocks(Mesh &mesh) {
    const size_t bs = (size_t)nvar * nvar;
    for (int l = 0; l < mesh.nline; ++l) {
        for (int k = 1; k < mesh.linesize[l]; ++k) {
            const int m = mesh.linepoint[l] + k;
            const int f = mesh.lineface[m];
            const int e = mesh.lines[m];
            const double *src = (e == mesh.fc[2 * f]) ? &jacO2[bs * f] : &jacO1[bs * f];
            std::memcpy(&A[bs * e], src, bs * sizeof(double));
        }
    }
}

void Jacobian::resizeBlockSize(int nvar_new) {
    if (nvar_new == nvar) return;
    // src/Jacobian.cxx:153-238, step for step in effect:
    //  - every array is resized in place (std::vector::resize keeps the old
    //    flat prefix and zero-fills what grows; the reference's zeroing loops
    //    iterate over copies and change nothing);
    //  - the diagonal of every D, Of1 and Of2 block is set to 10;
    //  - the leading min(old, new) square of each block is copied from the
    //    old block (old stride) into the new block (new stride);
    //  - rhs and U0 get a flat copy of their first nelem*min(old, new) values.
    // A and DinvC keep their old contents; the driver re-assembles A after a
    // resize (the reference leaves A at the old stride).
    printf("Resizing NVAR: old=%d new=%d ...", nvar, nvar_new);
    const int old = nvar, mn = std::min(nvar, nvar_new);
    const size_t nn = (size_t)nvar_new * nvar_new, oo = (size_t)old * old;
    auto reblock = [&](std::vector<double> &v, size_t nblk) {
        const std::vector<double> prev(v);
        v.resize(nn * nblk);
        for (size_t b = 0; b < nblk; ++b)
            for (int j = 0; j < nvar_new; ++j) v[b * nn + (size_t)nvar_new * j + j] = 10.0;
        for (size_t b = 0; b < nblk; ++b)
            for (int j = 0; j < mn; ++j)
                for (int i = 0; i < mn; ++i) v[b * nn + (size_t)nvar_new * j + i] = prev[b * oo + (size_t)old * j + i];
    };
    auto revec = [&](std::vector<double> &v) {
        const std::vector<double> prev(v);
        v.resize((size_t)nvar_new * nelem);
        for (size_t i = 0; i < (size_t)nelem * mn; ++i) v[i] = prev[i];
    };
    reblock(jacD, nelem);
    reblock(jacO1, nintface);
    reblock(jacO2, nintface);
    revec(rhs);
    revec(U0);
    nvar = nvar_new;
    const size_t bs = (size_t)nvar * nvar;
    jacDLU.resize(bs * nelem);
    A.resize(bs * nelem);
    DinvC.resize(bs * nelem);   // the device copy needs the new size
    U.assign((size_t)nvar * nelem, 0.0);
    dU.resize((size_t)nvar * nelem);
    res.resize((size_t)nvar * nelem);
    printf(" done!\n");
}

void Jacobian::setupDevice(Platform &gpu) {
    gpu_ = &gpu;
    o_jacDLU = gpu.malloc<double>(jacDLU.size());
    o_jacD = gpu.malloc<double>(jacD.size());
    o_jacO1 = gpu.malloc<double>(jacO1.size());
    o_jacO2 = gpu.malloc<double>(jacO2.size());
    o_rhs = gpu.malloc<double>(rhs.size());
    o_U0 = gpu.malloc<double>(U0.size());
    o_jacDinvC = gpu.malloc<double>(DinvC.size());
    o_U = gpu.malloc<double>(U.size());
    o_dU = gpu.malloc<double>(dU.size());
    o_res = gpu.malloc<double>(res.size());
    o_A = gpu.malloc<double>(A.size());
}

void Jacobian::toDevice() {
    auto up = [&](double *d, const std::vector<double> &h) {
        tbs_check(tbs_memcpy_h2d(gpu_->ctx, d, h.data(), h.size() * sizeof(double)), "Jacobian::toDevice");
    };
    up(o_jacDLU, jacDLU); up(o_jacD, jacD); up(o_jacO1, jacO1); up(o_jacO2, jacO2);
    up(o_rhs, rhs); up(o_U0, U0); up(o_U, U); up(o_res, res); up(o_dU, dU); up(o_A, A);
}

void Jacobian::fromDevice() {
    auto down = [&](std::vector<double> &h, const double *d) {
        tbs_check(tbs_memcpy_d2h(gpu_->ctx, h.data(), d, h.size() * sizeof(double)), "Jacobian::fromDevice");
    };
    down(jacDLU, o_jacDLU); down(jacD, o_jacD); down(jacO1, o_jacO1); down(jacO2, o_jacO2);
    down(rhs, o_rhs); down(U0, o_U0); down(U, o_U); down(dU, o_dU); down(res, o_res);
    down(DinvC, o_jacDinvC);
}
