/*
 * Jacobian.hxx -- block Jacobian values and LIJ vectors (the reference's
 * Jacobian class, include/Jacobian.hxx:43-98, src/Jacobian.cxx:11-290).
 * Layout: column-major nvar x nvar blocks, jacD by element, jacO1/jacO2 by
 * interior face, vectors by element.
 */
#ifndef TBS_JACOBIAN_HXX
#define TBS_JACOBIAN_HXX

#include <cstdint>
#include <string>
#include <vector>

#include "Mesh.hxx"
#include "Platform.hxx"

class Jacobian {
  public:
    size_t nbytes = 0;
    int nvar = 0, nelem = 0, nintface = 0;

    std::vector<double> jacDLU, jacD, jacO1, jacO2, rhs, res, U, U0, dU;
    std::vector<double> A, DinvC;

    double *o_jacDinvC = nullptr, *o_jacDLU = nullptr, *o_jacD = nullptr, *o_jacO1 = nullptr;
    double *o_jacO2 = nullptr, *o_rhs = nullptr, *o_res = nullptr, *o_U = nullptr;
    double *o_U0 = nullptr, *o_dU = nullptr, *o_A = nullptr;

    Jacobian() {}
    ~Jacobian();

    bool fromFile(int nlineelem, const std::string &path = "gpuline.jacobian.data.bin");
    /* seeded synthetic values: D = U[-1,1] + shift I, Of = U[-0.1,0.1], rhs = U[-1,1] */
    bool fromGenerator(const Mesh &mesh, uint64_t seed, double shift);
    bool toFile(const std::string &path) const;
    void assembleTriBlocks(Mesh &mesh);   /* src/Jacobian.cxx:76-117 (host) */
    void resizeBlockSize(int nvar_new);   /* src/Jacobian.cxx:153-238 semantics */
    void setupDevice(Platform &gpu);
    void toDevice();
    void fromDevice();

  private:
    void allocate(int nlineelem);
    Platform *gpu_ = nullptr;
};

#endif
