/*
 * Mesh.hxx -- line/face connectivity (the reference's Mesh class,
 * include/Mesh.hxx:27-67, src/Mesh.cxx:11-103).  Host std::vectors plus
 * device copies (raw HIP pointers owned by this object).  Indices are 0-based
 * after reading, as the reference's Fortran reader leaves them.
 */
#ifndef TBS_MESH_HXX
#define TBS_MESH_HXX

#include <string>
#include <vector>

#include "Platform.hxx"

class Mesh {
  public:
    size_t nbytes = 0;
    int nvar = 0, nelem = 0, nintface = 0, eftot = 0;
    int nline = 0, nlineelem = 0, max_line_nelem = 0;

    std::vector<int> epoint, ef, fc, lines, linesize, lineface, linepoint;

    int *o_epoint = nullptr, *o_ef = nullptr, *o_fc = nullptr, *o_lines = nullptr;
    int *o_linesize = nullptr, *o_lineface = nullptr, *o_linepoint = nullptr;

    Mesh() {}
    ~Mesh();

    /* gpuline.mesh.data.bin (cwd by default), size-checked; false on error */
    bool fromFile(const std::string &path = "gpuline.mesh.data.bin");
    /* seeded-workload topology: NL lines of NK elements */
    bool fromStructured(int NL, int NK, int nvar);
    bool toFile(const std::string &path) const;
    void setupDevice(Platform &gpu);
    void toDevice();
    void fromDevice();
    tbs_mesh_host host() const;

  private:
    void finishRead();
    Platform *gpu_ = nullptr;
};

#endif
