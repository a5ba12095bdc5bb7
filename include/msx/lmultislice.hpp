// msx/lmultislice.hpp -- header-only C++ mirror of the reference's LMultiSlice
// call surface over the msx C ABI (include/msx.h):
//
//   Microscope        ImageSimulation/LMultiSlice/Microscope.h:3-34
//   ModelPotential    ImageSimulation/LMultiSlice/ModelPotential.h:27-51
//   ModelSimulated    ImageSimulation/LMultiSlice/ModelSimulated.h:7-28
//   CPotentialBuilder ImageSimulation/PotentialBuilder/PotentialBuilder.h:14-22
//   AModel::Model-compatible atom container (amodel/include/Model.h:20-174,
//   BaseStructures.h:50-135) with the .coo reader (ModelCoo.cpp:20-90) and an
//   ImageSpace::Image-compatible result (ImageSimulation/Image/Image.h:29-88)
//   for callers without those DLLs;
//   Dispatcher        ImageSimulation/LMultiSlice/Dispatcher.cu:25-242 (the
//                     XML command and the run: potential, _pt.mrc, image,
//                     .mrc), driven by the `lmultislice` executable.
//
// The classes live in namespace msx::lms.  Defining MSX_LMS_GLOBAL_NAMES before
// including this header also exports ModelPotential, ModelSimulated,
// Microscope and CPotentialBuilder into the global namespace, as the reference
// declares them.  Member names, argument meanings and the 0 / -1 return
// convention follow the reference; all computation goes to libmsx.so on the
// GPU (one process-wide context on device 0 unless set_context() is called).
//
// Physics differences from LMultiSlice, by design (its k^4 propagator,
// per-slice fftshift and 1/nx normalisation are not reproduced): the slice
// loop is Kirkland's (transmission with the 2/3 band limit, Fresnel propagator
// exp(-i pi lambda dz k^2), forward FFT unscaled / inverse scaled by
// 1/(nx ny)), and the lens is the Kirkland CTF exp(-i chi) with a hard
// objective aperture.
#ifndef MSX_LMULTISLICE_HPP
#define MSX_LMULTISLICE_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <initializer_list>
#include <iostream>
#include <iterator>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "../msx.h"

namespace msx {
namespace lms {

// ----------------------------------------------------------------- context
inline msx_ctx& context_slot() {
    static msx_ctx c = nullptr;
    return c;
}
// Process-wide context (device 0) created on first use.
inline msx_ctx context() {
    msx_ctx& c = context_slot();
    if (!c && msx_create(0, &c) != MSX_OK) c = nullptr;
    return c;
}
inline void set_context(msx_ctx c) { context_slot() = c; }

// ----------------------------------------------------------------- atoms
namespace amodel {
struct XYZ {
    float x = 0, y = 0, z = 0;
};
struct Record {  // BaseStructures.h:120-128
    char Atom[4] = {0, 0, 0, 0};
    float oxygenation = 0;
    XYZ xsCoordinate;  // fractional coordinates
    float R_e = 0, occupation = 1, absorbtion = 0;
};
struct Cortege {  // BaseStructures.h:129-135
    bool isInclude = true;
    Record element;
};

inline const char* const* element_names() {
    static const char* table[] = {
        "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar",
        "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr",
        "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe",
        "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf",
        "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th",
        "Pa", "U",  "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es", "Fm", "Md", "No", "Lr"};
    return table;
}
// Z from the element symbol (Model::getNumberByName, Model.h:62-81).
inline int number_by_name(const char* name) {
    for (int i = 0; i < 103; i++)
        if (std::strcmp(name, element_names()[i]) == 0) return i + 1;
    return -1;
}
// Symbol from Z (Model::getNameByNumber, Model.h:82-105); "" outside 1..103.
inline const char* name_by_number(int z) { return (z >= 1 && z <= 103) ? element_names()[z - 1] : ""; }

// A number of a .coo file (Model::StringToDouble, Model.h:106-131): "p/q"
// is a fraction, a decimal comma reads as a point.  (The reference turns
// character 1 into the point, which is the comma only for one-digit integer
// parts; here the comma itself is replaced.)
inline double coo_number(const char* s) {
    std::string t(s ? s : "");
    const size_t sl = t.find('/');
    if (sl != std::string::npos) return std::atof(t.substr(0, sl).c_str()) / std::atof(t.substr(sl + 1).c_str());
    for (char& ch : t)
        if (ch == ',') ch = '.';
    return std::atof(t.c_str());
}

// Minimal container with the accessors ModelPotential uses on AModel::Model.
class Model {
public:
    Model() = default;
    Model(float a, float b, float c) : a_(a), b_(b), c_(c) {}
    float getA() const { return a_; }
    float getB() const { return b_; }
    float getC() const { return c_; }
    float getAlpha() const { return alpha_; }
    float getBeta() const { return beta_; }
    float getGamma() const { return gamma_; }
    size_t getNumberAtoms() const { return cells_.size(); }
    Cortege* getTableCell() { return cells_.data(); }
    int getNumberByName(const char* name) const { return number_by_name(name); }
    void setA(float v) { a_ = v; }
    void setB(float v) { b_ = v; }
    void setC(float v) { c_ = v; }
    void setAlpha(float v) { alpha_ = v; }
    void setBeta(float v) { beta_ = v; }
    void setGamma(float v) { gamma_ = v; }
    void add(const char* symbol, float fx, float fy, float fz, float occ = 1.f) {
        Cortege c;
        std::strncpy(c.element.Atom, symbol, 3);
        c.element.xsCoordinate = XYZ{fx, fy, fz};
        c.element.occupation = occ;
        cells_.push_back(c);
    }
    // ModelCoo::read (ModelCoo.cpp:20-90): a comment line; "a b c alpha beta
    // gamma n" (decimal commas allowed); then n lines "x y z Z absorption
    // oxygenation occupation" with fractional coordinates.  0 / -1.
    int read(const char* fileName) {
        std::ifstream fin(fileName);
        if (!fin.is_open()) {
            std::cerr << "File with name " << fileName << " doesn't exists!" << std::endl;
            return -1;
        }
        std::string line;
        if (!std::getline(fin, line) || !std::getline(fin, line)) return -1;
        for (char& ch : line)
            if (ch == ',') ch = '.';
        std::istringstream head(line);
        double v[6];
        long long n = -1;
        for (double& x : v)
            if (!(head >> x)) return -1;
        if (!(head >> n) || n < 0) return -1;
        a_ = (float)v[0], b_ = (float)v[1], c_ = (float)v[2];
        alpha_ = (float)v[3], beta_ = (float)v[4], gamma_ = (float)v[5];
        cells_.clear();
        cells_.reserve((size_t)n);
        for (long long i = 0; i < n; i++) {
            if (!std::getline(fin, line)) return -1;
            std::istringstream row(line);
            std::string tok[7];
            for (auto& t : tok)
                if (!(row >> t)) return -1;
            Cortege c;
            c.element.xsCoordinate = XYZ{(float)coo_number(tok[0].c_str()), (float)coo_number(tok[1].c_str()),
                                         (float)coo_number(tok[2].c_str())};
            std::strncpy(c.element.Atom, name_by_number(std::atoi(tok[3].c_str())), 3);
            c.element.absorbtion = (float)coo_number(tok[4].c_str());
            c.element.oxygenation = (float)coo_number(tok[5].c_str());
            c.element.occupation = (float)coo_number(tok[6].c_str());
            c.element.R_e = 0.2f;
            cells_.push_back(c);
        }
        return 0;
    }

private:
    float a_ = 0, b_ = 0, c_ = 0, alpha_ = 90, beta_ = 90, gamma_ = 90;
    std::vector<Cortege> cells_;
};
}  // namespace amodel

// ----------------------------------------------------------------- image
// Interleaved image with the fields of ImageSpace::Image that callers read
// (nChannels, depth in bytes per value, width, height, thickness, imageData).
struct Image {
    int nChannels = 1, depth = 8, width = 0, height = 0, thickness = 1;
    std::vector<double> store;
    char* imageData = nullptr;
    Image() = default;
    Image(size_t nx, size_t ny, size_t nz, size_t depth_, size_t nCh)
        : nChannels((int)nCh), depth((int)depth_), width((int)nx), height((int)ny), thickness((int)nz),
          store(nx * ny * nz * nCh, 0.0) {
        imageData = reinterpret_cast<char*>(store.data());
    }
    double* data() { return store.data(); }
    // |psi| of a 2-channel complex image (Image::getModule, Image.cu:234-257)
    Image getModule() const {
        Image m((size_t)width, (size_t)height, (size_t)thickness, 8, 1);
        for (size_t i = 0; i < m.store.size(); i++)
            m.store[i] = std::hypot(store[nChannels * i], nChannels > 1 ? store[nChannels * i + 1] : 0.0);
        return m;
    }
    // <filename>.mrc in an MRC mode (Image::saveMRC, Image.cu:138-163 with
    // MRC::save, MRC.cpp:206-345), cell from the model; defined below.
    template <class ModelT>
    int saveMRC(const char* filename, const ModelT* model, int w, int h, int thick, int mode = 2) const;
};

// MRC data modes (ImageSpace::mrcMode, otherEnums.h:21-29) and their
// channels per pixel (MRC::nChannels, MRC.h:108-113).
enum mrcMode { mrc_BYTE = 0, mrc_SHORT = 1, mrc_FLOAT = 2, mrc_SHORT2 = 3, mrc_FLOAT2 = 4, mrc_INT = 6, mrc_CHAR3 = 16 };
inline int mrc_channels(int mode) { return (mode == 3 || mode == 4) ? 2 : (mode == 16 ? 3 : 1); }

namespace detail {
// double -> integer element of an MRC mode: truncation toward zero as the
// reference's C conversions (MRC.cpp:310-333), saturated at the type's range
// (out of range the reference's conversion is undefined) and NaN -> 0
template <typename T> T mrc_int(double x) {
    if (!(x == x)) return 0;
    const double lo = (double)std::numeric_limits<T>::min(), hi = (double)std::numeric_limits<T>::max();
    return (T)std::trunc(std::min(std::max(x, lo), hi));
}
}  // namespace detail

// MRC2000 writer (MRC::save, MRC.cpp:206-345): a 1024-byte header, then
// nx*ny*nz pixels of mrc_channels(mode) values each, taken from v in order
// (v holds nx*ny*nz*mrc_channels(mode) doubles) and stored as mode says:
// 0 and 16 unsigned bytes, 1 and 3 signed shorts, 2 and 4 floats, 6 unsigned
// shorts.  The reference writes nx*ny*nz*channels*4 bytes whatever the mode
// (past the end of its converted buffer for modes 0, 1, 3 and 6, and from a
// null buffer for mode 16, whose case lacks a break, MRC.cpp:304-307); here
// each mode writes exactly its values.  Header: sizes, mode, cell lengths and
// angles (alpha, beta, gamma; 90 when not given), axis order 1 2 3, the
// minimum, maximum, mean and RMS deviation of the stored values, "MAP " and a
// little-endian machine stamp.  Returns 0, or -1 for an unknown mode or a
// failed write.
inline int write_mrc(const char* path, const double* v, int nx, int ny, int nz, float xlen, float ylen, float zlen,
                     int mode = mrc_FLOAT, float alpha = 90.f, float beta = 90.f, float gamma = 90.f) {
    if (mode != 0 && mode != 1 && mode != 2 && mode != 3 && mode != 4 && mode != 6 && mode != 16) return -1;
    const size_t n = (size_t)nx * ny * nz * mrc_channels(mode);
    // the stored values, as doubles for the header statistics
    std::vector<double> q(n);
    std::vector<unsigned char> data;
    auto put = [&](const void* p, size_t bytes) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        data.insert(data.end(), c, c + bytes);
    };
    data.reserve(n * 4);
    for (size_t i = 0; i < n; i++) {
        if (mode == 0 || mode == 16) {
            const unsigned char x = detail::mrc_int<unsigned char>(v[i]);
            put(&x, 1);
            q[i] = x;
        } else if (mode == 1 || mode == 3) {
            const int16_t x = detail::mrc_int<int16_t>(v[i]);
            put(&x, 2);
            q[i] = x;
        } else if (mode == 6) {
            const uint16_t x = detail::mrc_int<uint16_t>(v[i]);
            put(&x, 2);
            q[i] = x;
        } else {
            const float x = (float)v[i];
            put(&x, 4);
            q[i] = x;
        }
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return -1;
    unsigned char h[1024];
    std::memset(h, 0, sizeof h);
    auto put_i = [&](int word, int32_t x) { std::memcpy(h + 4 * word, &x, 4); };
    auto put_f = [&](int word, float x) { std::memcpy(h + 4 * word, &x, 4); };
    double mn = n ? q[0] : 0, mx = mn, sum = 0;
    for (size_t i = 0; i < n; i++) {
        mn = std::min(mn, q[i]);
        mx = std::max(mx, q[i]);
        sum += q[i];
    }
    const double mean = n ? sum / n : 0;
    double var = 0;
    for (size_t i = 0; i < n; i++) var += (q[i] - mean) * (q[i] - mean);
    put_i(0, nx); put_i(1, ny); put_i(2, nz); put_i(3, mode);
    put_i(7, nx); put_i(8, ny); put_i(9, nz);
    put_f(10, xlen); put_f(11, ylen); put_f(12, zlen);
    put_f(13, alpha); put_f(14, beta); put_f(15, gamma);
    put_i(16, 1); put_i(17, 2); put_i(18, 3);
    put_f(19, (float)mn); put_f(20, (float)mx); put_f(21, (float)mean);
    std::memcpy(h + 208, "MAP ", 4);
    h[212] = 0x44; h[213] = 0x44;  // little-endian machine stamp
    put_f(54, (float)std::sqrt(n ? var / n : 0));
    bool ok = std::fwrite(h, 1, 1024, f) == 1024;
    ok = ok && (data.empty() || std::fwrite(data.data(), 1, data.size(), f) == data.size());
    return (std::fclose(f) == 0 && ok) ? 0 : -1;
}

template <class ModelT>
int Image::saveMRC(const char* filename, const ModelT* model, int w, int h, int thick, int mode) const {
    // MRC::save reads the image's interleaved values in order, mrc_channels(mode)
    // per pixel (MRC.cpp:308-333 over pData = imageData, Image.cu:141-158): for
    // an image with as many channels as the mode, pixel by pixel
    const size_t n = (size_t)w * h * thick * mrc_channels(mode);
    if (w < 0 || h < 0 || thick < 0 || n > (size_t)width * height * thickness * nChannels) return -1;
    const std::vector<double> v(store.begin(), store.begin() + (long)n);
    const std::string p = std::string(filename) + ".mrc";
    return write_mrc(p.c_str(), v.data(), w, h, thick, model ? model->getA() : 0.f, model ? model->getB() : 0.f,
                     model ? model->getC() : 0.f, mode, model ? model->getAlpha() : 90.f,
                     model ? model->getBeta() : 90.f, model ? model->getGamma() : 90.f);
}

// ----------------------------------------------------------------- Microscope
// Microscope.cpp:8-86.  Units as there: keV, Cs [mm], aperture [mrad]
// (objective aperture semi-angle), defocus [nm].
class Microscope {
public:
    Microscope() = default;
    Microscope(double keV, double Cs, double aperture, double defocus)
        : keV_(keV), Cs_(Cs), aperture_(aperture), defocus_(defocus) {}
    double getKeV() const { return keV_; }
    double getCs() const { return Cs_; }
    double getAperture() const { return aperture_; }
    double getDefocus() const { return defocus_; }
    double getLambda() const { return msx_wavelength(keV_); }
    double getSigma() const { return msx_sigma(keV_); }
    void setKeV(double v) { keV_ = v; }
    void setCs(double v) { Cs_ = v; }
    void setAperture(double v) { aperture_ = v; }
    void setDefocus(double v) { defocus_ = v; }
    // Microscope::Es (Microscope.cpp:17-21): the envelope
    // exp(-(pi aperture / (1000 lambda))^2 (Cs lambda^3 k^3 + defocus lambda k)^2)
    // for a spatial frequency k, with Cs and defocus taken as stored (no unit
    // conversion, as there).
    double Es(double k) const {
        const double l = getLambda();
        const double a = M_PI * aperture_ / (l * 1000.0);
        const double b = Cs_ * l * l * l * k * k * k + defocus_ * l * k;
        return std::exp(-(a * a) * (b * b));
    }
    // Microscope::alpha (Microscope.cpp:23-25): pi lambda k^2 (0.5 Cs lambda^2 k^2
    // - 10 defocus) for a spatial frequency k, Cs as stored.
    double alpha(double k) const {
        const double l = getLambda();
        return M_PI * l * k * k * (0.5 * Cs_ * l * l * k * k - defocus_ * 10.0);
    }
    // Kirkland's aberration phase chi = pi lambda k2 (0.5 Cs lambda^2 k2 - df) for
    // |k|^2 = k2 (1/A^2), with Cs converted from mm and defocus from nm: the
    // phase of the objective lens that imageCalculation applies
    // (slicelib.cpp:258-318 without the higher aberrations).
    double chi_k2(double k2) const {
        const double l = getLambda(), cs = Cs_ * 1e7, df = defocus_ * 10.0;
        return M_PI * l * k2 * (0.5 * cs * l * l * k2 - df);
    }

private:
    double keV_ = 200, Cs_ = 0, aperture_ = 0, defocus_ = 0;
};

// ----------------------------------------------------------------- ModelPotential
// ModelPotential.h:27-51.  `potential` is a host double[nz][ny][nx] (x
// fastest, the reference's layout) filled by calculatePotentialGrid().
// The default reproduces the LMultiSlice kernel's formula and table
// (ModelPotential.cu:232-254); set_formula(MSX_FORMULA_KIRKLAND) selects
// Kirkland's vzatom.
template <class ModelT = amodel::Model> class ModelPotentialT {
public:
    // ModelPotential(void) (ModelPotential.h:29): no model and an empty grid;
    // calculatePotentialGrid() then returns -1.
    ModelPotentialT() : potential(nullptr), model_(nullptr), nx_(0), ny_(0), nz_(0), dpa_(0), radius_(0) {}
    ModelPotentialT(ModelT* model, size_t nx, size_t ny, size_t nz, double dpa, double radius)
        : potential(nullptr), model_(model), nx_(nx), ny_(ny), nz_(nz), dpa_(dpa), radius_(radius) {
        store_.assign(nx * ny * nz, 0.0);
        potential = store_.data();
    }
    int calculatePotentialGrid() {
        if (!model_ || nx_ * ny_ * nz_ == 0) return -1;
        msx_ctx c = context();
        if (!c) return -1;
        const size_t n = model_->getNumberAtoms();
        auto* cells = model_->getTableCell();
        const double a = model_->getA(), b = model_->getB(), cc = model_->getC();
        std::vector<msx_atom> atoms;
        atoms.reserve(n);
        for (size_t i = 0; i < n; i++) {
            const auto& e = cells[i].element;
            msx_atom at;
            at.Z = model_->getNumberByName(e.Atom);
            if (at.Z < 1) return -1;
            at.occ = use_occupancy_ ? e.occupation : 1.0f;
            at.x = e.xsCoordinate.x * a;
            at.y = e.xsCoordinate.y * b;
            at.z = (double)e.xsCoordinate.z * cc;
            atoms.push_back(at);
        }
        std::vector<double> v(nx_ * ny_ * nz_);  // [kz][ix][iy] from the engine
        if (msx_potential(c, atoms.data(), (int64_t)atoms.size(), a, b, cc, (int)nx_, (int)ny_, (int)nz_, radius_,
                          table_, formula_, v.data(), MSX_F64, MSX_HOST) != MSX_OK)
            return -1;
        for (size_t kz = 0; kz < nz_; kz++)
            for (size_t ix = 0; ix < nx_; ix++)
                for (size_t iy = 0; iy < ny_; iy++)
                    potential[(kz * ny_ + iy) * nx_ + ix] = v[(kz * nx_ + ix) * ny_ + iy];
        return 0;
    }
    // <filename>_pt.mrc, as ModelPotential::savePotential (ModelPotential.cu:294-304)
    int savePotential(const char* filename) {
        std::string p = std::string(filename) + "_pt.mrc";
        return write_mrc(p.c_str(), potential, (int)nx_, (int)ny_, (int)nz_, model_ ? model_->getA() : 0,
                         model_ ? model_->getB() : 0, model_ ? model_->getC() : 0);
    }
    ModelT* getModel() { return model_; }
    size_t getNx() const { return nx_; }
    size_t getNy() const { return ny_; }
    size_t getNz() const { return nz_; }
    double getDpa() const { return dpa_; }
    void set_formula(int f) { formula_ = f; }
    void set_table(int t) { table_ = t; }
    void set_use_occupancy(bool u) { use_occupancy_ = u; }

    double* potential;

private:
    std::vector<double> store_;
    ModelT* model_;
    size_t nx_, ny_, nz_;
    double dpa_, radius_;
    int table_ = MSX_TABLE_LMULTISLICE, formula_ = MSX_FORMULA_LMULTISLICE;
    bool use_occupancy_ = false;
};
using ModelPotential = ModelPotentialT<amodel::Model>;

// ----------------------------------------------------------------- ModelSimulated
// ModelSimulated.h:7-28.  imageCalculation(result, microscope): plane wave
// through the nz slices of the potential (thickness c/nz each), then the
// objective lens; result receives the image-plane wave as interleaved complex
// doubles (nChannels 2), like Image::copyFFTtoImage (Image.h:62-73).
template <class MP = ModelPotential> class ModelSimulatedT {
public:
    // ModelSimulated(void) (ModelSimulated.h:9): no potential; imageCalculation
    // then returns -1.
    ModelSimulatedT() : mp_(nullptr), nx_(0), ny_(0), nz_(0), dpa_(0) {}
    ModelSimulatedT(MP* mp, size_t nx, size_t ny, size_t nz, double dpa)
        : mp_(mp), nx_(nx), ny_(ny), nz_(nz), dpa_(dpa) {}
    template <class ImageT> int imageCalculation(ImageT* result, const Microscope* mic) {
        if (!mp_ || !mp_->getModel() || !result || !mic || nx_ * ny_ * nz_ == 0) return -1;
        msx_ctx c = context();
        if (!c) return -1;
        auto* model = mp_->getModel();
        const double a = model->getA(), b = model->getB(), cc = model->getC();
        const size_t n = nx_ * ny_;
        // potential [kz][iy][ix] -> engine layout [kz][ix][iy]
        std::vector<float> V(n * nz_);
        for (size_t kz = 0; kz < nz_; kz++)
            for (size_t iy = 0; iy < ny_; iy++)
                for (size_t ix = 0; ix < nx_; ix++)
                    V[(kz * nx_ + ix) * ny_ + iy] = (float)mp_->potential[(kz * ny_ + iy) * nx_ + ix];
        std::vector<float> t(2 * n * nz_);
        const double kev = mic->getKeV();
        if (msx_transmission(c, V.data(), (int)nz_, (int)nx_, (int)ny_, a, b, kev, 2.0 / 3.0, t.data(), MSX_F32,
                             MSX_HOST) != MSX_OK)
            return -1;
        std::vector<double> cz(nz_, cc / nz_);
        std::vector<int32_t> layer(nz_);
        for (size_t k = 0; k < nz_; k++) layer[k] = (int32_t)k;
        msx_stack st{};
        st.nx = (int32_t)nx_;
        st.ny = (int32_t)ny_;
        st.ax = a;
        st.by = b;
        st.kev = kev;
        st.bw = 2.0 / 3.0;
        st.nlayer = (int32_t)nz_;
        st.cz = cz.data();
        st.nslice = (int32_t)nz_;
        st.layer = layer.data();
        std::vector<float> one(2 * n, 0.f), psi(2 * n);
        for (size_t i = 0; i < n; i++) one[2 * i] = 1.f;
        if (msx_multislice(c, &st, t.data(), one.data(), psi.data(), 1, MSX_F32, MSX_HOST) != MSX_OK) return -1;
        exit_wave_.assign(psi.begin(), psi.end());
        std::vector<float> img(2 * n);
        if (msx_ctf_image(c, psi.data(), 1, (int)nx_, (int)ny_, a, b, kev, mic->getCs() * 1e7,
                          mic->getDefocus() * 10.0, mic->getAperture() * 1e-3, img.data(), nullptr, MSX_F32,
                          MSX_HOST) != MSX_OK)
            return -1;
        // engine (ix,iy) -> image row iy, column ix, interleaved (re, im)
        double* out = reinterpret_cast<double*>(result->imageData);
        const int nch = result->nChannels;
        for (size_t iy = 0; iy < ny_; iy++)
            for (size_t ix = 0; ix < nx_; ix++) {
                out[nch * (iy * nx_ + ix) + 0] = img[2 * (ix * ny_ + iy)];
                if (nch > 1) out[nch * (iy * nx_ + ix) + 1] = img[2 * (ix * ny_ + iy) + 1];
            }
        return 0;
    }
    // exit wave of the last imageCalculation (engine layout, complex float)
    const std::vector<float>& exitWave() const { return exit_wave_; }

private:
    MP* mp_;
    size_t nx_, ny_, nz_;
    double dpa_;
    std::vector<float> exit_wave_;
};
using ModelSimulated = ModelSimulatedT<ModelPotential>;

// ----------------------------------------------------------------- PotentialBuilder
// PotentialBuilder.h:14-22: the reference DLL is an empty skeleton that
// exports a class with an empty constructor, the variable nPotentialBuilder
// (0) and the function fnPotentialBuilder() (returns 42).  The same three
// symbols are defined once in libmsx.so (msx_formats.hip), so a caller links
// them from the library as it linked them from PotentialBuilder.dll.
class MSX_API CPotentialBuilder {
public:
    CPotentialBuilder(void);
};
extern MSX_API int nPotentialBuilder;
MSX_API int fnPotentialBuilder(void);

// ----------------------------------------------------------------- Dispatcher
// The run command of an LMultiSlice XML file (Dispatcher.h, Dispatcher.cu:
// 25-159): <action> <io> fileNameInput, fileNameOutput </io> <image> nx, ny,
// dpa </image> <slicing> radiuc, numberslices </slicing> <microscope>
// aperture [mrad], cs [mm], defocus [nm], keV </microscope> </action>.
struct Command {
    std::string fileNameInput, fileNameOutput;
    int nx = 0, ny = 0, numberSlices = 0;
    float dpa = 0, radiuc = 0, aperture = 0, cs = 0, defocus = 0, keV = 0;
};

// Text of the element reached by the path of child names (the first match
// at each level), trimmed; "" when absent.  Enough XML for the command files:
// comments and declarations are skipped, attributes are ignored.
inline std::string xml_child_value(const std::string& doc, std::initializer_list<const char*> path) {
    size_t lo = 0, hi = doc.size();
    for (const char* name : path) {
        const std::string open = std::string("<") + name, close = std::string("</") + name + ">";
        size_t at = lo, start = std::string::npos;
        while ((at = doc.find(open, at)) != std::string::npos && at < hi) {
            const char next = at + open.size() < doc.size() ? doc[at + open.size()] : '\0';
            if (next == '>' || next == ' ' || next == '\t' || next == '\n' || next == '\r') {
                start = doc.find('>', at);
                break;
            }
            at += open.size();
        }
        if (start == std::string::npos || start >= hi) return "";
        const size_t end = doc.find(close, start);
        if (end == std::string::npos || end > hi) return "";
        lo = start + 1;
        hi = end;
    }
    std::string v = doc.substr(lo, hi - lo);
    const size_t b = v.find_first_not_of(" \t\r\n"), e = v.find_last_not_of(" \t\r\n");
    return b == std::string::npos ? "" : v.substr(b, e - b + 1);
}

// Dispatcher::parseCommand: 0, or -1 with the reference's message when the
// file cannot be read or a field is empty.
inline int parse_command(const char* fileNameXML, Command& command) {
    std::ifstream fin(fileNameXML, std::ios::binary);
    if (!fin.is_open()) {
        std::cerr << "XML File with name [" << fileNameXML << "] doesn't exist." << std::endl;
        return -1;
    }
    std::string doc((std::istreambuf_iterator<char>(fin)), std::istreambuf_iterator<char>());
    // drop comments and declarations
    for (size_t a; (a = doc.find("<!--")) != std::string::npos;) {
        const size_t b = doc.find("-->", a + 4);
        doc.erase(a, b == std::string::npos ? std::string::npos : b + 3 - a);
    }
    for (size_t a; (a = doc.find("<?")) != std::string::npos;) {
        const size_t b = doc.find("?>", a + 2);
        doc.erase(a, b == std::string::npos ? std::string::npos : b + 2 - a);
    }
    struct Field {
        const char* group;
        const char* name;
        const char* label;
    };
    auto get = [&](const Field& f, std::string& out) -> int {
        out = xml_child_value(doc, {"action", f.group, f.name});
        if (out.empty()) {
            std::cerr << "Empty " << f.label << " field!" << std::endl;
            return -1;
        }
        return 0;
    };
    std::string v;
    if (get({"io", "fileNameInput", "file input"}, command.fileNameInput)) return -1;
    if (get({"io", "fileNameOutput", "file output"}, command.fileNameOutput)) return -1;
    if (get({"image", "nx", "nx"}, v)) return -1;
    command.nx = std::atoi(v.c_str());
    if (get({"image", "ny", "ny"}, v)) return -1;
    command.ny = std::atoi(v.c_str());
    if (get({"image", "dpa", "dpa"}, v)) return -1;
    command.dpa = (float)std::atof(v.c_str());
    if (get({"slicing", "radiuc", "radiuc"}, v)) return -1;
    command.radiuc = (float)std::atof(v.c_str());
    if (get({"slicing", "numberslices", "numberslices"}, v)) return -1;
    command.numberSlices = std::atoi(v.c_str());
    if (get({"microscope", "aperture", "aperture"}, v)) return -1;
    command.aperture = (float)std::atof(v.c_str());
    if (get({"microscope", "cs", "cs"}, v)) return -1;
    command.cs = (float)std::atof(v.c_str());
    if (get({"microscope", "defocus", "defocus"}, v)) return -1;
    command.defocus = (float)std::atof(v.c_str());
    if (get({"microscope", "keV", "keV"}, v)) return -1;
    command.keV = (float)std::atof(v.c_str());
    return 0;
}

// Dispatcher::Run (Dispatcher.cu:161-242): read the command and the model
// (.coo, getModelType), build the potential and write <output>_pt.mrc, run
// the plane wave through the slices and the objective lens, write |psi| of
// the image wave to <output>.mrc.  0 / -1.
inline int dispatcher_run(const char* fileNameXML) {
    Command command;
    if (parse_command(fileNameXML, command) == -1) return -1;
    const std::string in = command.fileNameInput;
    if (in.size() < 4 || in.compare(in.size() - 4, 4, ".coo") != 0) {
        std::cout << "Can not read file " << in << "!!! (only .coo models are supported)" << std::endl;
        return -1;
    }
    amodel::Model model;
    if (model.read(in.c_str()) == -1) {
        std::cout << "Can not read file " << in << "!!!" << std::endl;
        return -1;
    }
    std::cout << "Read file model [" << in << "] successful." << std::endl;
    std::cout << "Image size = " << command.nx << "x" << command.ny << std::endl;
    std::cout << "Number slides = " << command.numberSlices << std::endl;
    std::cout << "dpa = " << command.dpa << std::endl;
    ModelPotential potential(&model, (size_t)command.nx, (size_t)command.ny, (size_t)command.numberSlices,
                             command.dpa, command.radiuc);
    if (potential.calculatePotentialGrid() == -1) return -1;
    if (potential.savePotential(command.fileNameOutput.c_str()) == -1) return -1;
    ModelSimulated simulated(&potential, (size_t)command.nx, (size_t)command.ny, (size_t)command.numberSlices,
                             command.dpa);
    Microscope microscope(command.keV, command.cs, command.aperture, command.defocus);
    Image result((size_t)command.nx, (size_t)command.ny, 1, sizeof(double), 2);
    if (simulated.imageCalculation(&result, &microscope) == -1) return -1;
    const Image module = result.getModule();
    if (module.saveMRC(command.fileNameOutput.c_str(), &model, command.nx, command.ny, 1) == -1) return -1;
    std::cout << "Calculation for [" << fileNameXML << "] finished successful." << std::endl << std::endl;
    return 0;
}

}  // namespace lms
}  // namespace msx

#ifdef MSX_LMS_GLOBAL_NAMES
using msx::lms::CPotentialBuilder;
using msx::lms::fnPotentialBuilder;
using msx::lms::nPotentialBuilder;
using msx::lms::Microscope;
using msx::lms::ModelPotential;
using msx::lms::ModelSimulated;
#endif

#endif  // MSX_LMULTISLICE_HPP
