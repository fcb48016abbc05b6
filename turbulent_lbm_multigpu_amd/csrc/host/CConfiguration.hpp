// CConfiguration.hpp -- the conf.xml surface (src/CConfiguration.hpp:21-161,
// conf.xml:1-56) with the same tags, including the `domian-length` spelling,
// parsed by a small built-in XML reader (elements and text only; tinyxml2 is
// not available).  An optional <lattice> block gives lattice-unit values that
// bypass the physical parametrization:
//   <lattice> <tau/> <lid-velocity/> <body-force><x/><y/><z/></body-force>
//             <smagorinsky-cs/> <storage>f32|f64|f16|bf16</storage>
//             <periodic><x/><y/><z/></periodic> <forcing>reference|body</forcing>
//             <kernel-variant/> <exchange-mode/> <fused-x/> <alpha-work-group/>
//             <perturbation><amplitude/><seed/></perturbation>
//   </lattice>
#pragma once

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <list>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "CLbmSolver.hpp"
#include "CVector.hpp"

namespace lbm {

struct XmlNode {
    std::string name, text;
    std::vector<std::unique_ptr<XmlNode>> children;
    const XmlNode *child(const std::string &n) const {
        for (auto &c : children)
            if (c->name == n) return c.get();
        return nullptr;
    }
};

// Elements, text, comments, <?...?> and <!...> declarations; attributes are
// skipped.  Throws std::runtime_error on malformed input.
class XmlReader {
    const std::string &s_;
    size_t p_ = 0;
    [[noreturn]] void bad(const char *what) {
        throw std::runtime_error(std::string("XML parse error: ") + what + " at offset " + std::to_string(p_));
    }
    void skip_misc() {
        for (;;) {
            while (p_ < s_.size() && isspace((unsigned char)s_[p_])) ++p_;
            if (s_.compare(p_, 4, "<!--") == 0) {
                size_t e = s_.find("-->", p_ + 4);
                if (e == std::string::npos) bad("unterminated comment");
                p_ = e + 3;
            } else if (s_.compare(p_, 2, "<?") == 0) {
                size_t e = s_.find("?>", p_ + 2);
                if (e == std::string::npos) bad("unterminated declaration");
                p_ = e + 2;
            } else if (s_.compare(p_, 2, "<!") == 0) {
                size_t e = s_.find('>', p_ + 2);
                if (e == std::string::npos) bad("unterminated declaration");
                p_ = e + 1;
            } else {
                return;
            }
        }
    }
    std::unique_ptr<XmlNode> element() {
        if (p_ >= s_.size() || s_[p_] != '<') bad("expected '<'");
        ++p_;
        size_t b = p_;
        while (p_ < s_.size() && !isspace((unsigned char)s_[p_]) && s_[p_] != '>' && s_[p_] != '/') ++p_;
        auto n = std::make_unique<XmlNode>();
        n->name = s_.substr(b, p_ - b);
        if (n->name.empty()) bad("empty tag name");
        size_t close = s_.find('>', p_);
        if (close == std::string::npos) bad("unterminated tag");
        if (s_[close - 1] == '/') {
            p_ = close + 1;
            return n;
        }
        p_ = close + 1;
        for (;;) {
            size_t t = p_;
            while (p_ < s_.size() && s_[p_] != '<') ++p_;
            if (p_ >= s_.size()) bad("unterminated element");
            n->text += s_.substr(t, p_ - t);
            if (s_.compare(p_, 4, "<!--") == 0) {
                size_t e = s_.find("-->", p_ + 4);
                if (e == std::string::npos) bad("unterminated comment");
                p_ = e + 3;
                continue;
            }
            if (s_.compare(p_, 2, "</") == 0) {
                size_t e = s_.find('>', p_);
                if (e == std::string::npos) bad("unterminated end tag");
                std::string endname = s_.substr(p_ + 2, e - p_ - 2);
                while (!endname.empty() && isspace((unsigned char)endname.back())) endname.pop_back();
                if (endname != n->name) bad("mismatched end tag");
                p_ = e + 1;
                break;
            }
            n->children.push_back(element());
        }
        size_t a = n->text.find_first_not_of(" \t\r\n"), z = n->text.find_last_not_of(" \t\r\n");
        n->text = a == std::string::npos ? std::string() : n->text.substr(a, z - a + 1);
        return n;
    }

public:
    explicit XmlReader(const std::string &s) : s_(s) {}
    std::unique_ptr<XmlNode> parse() {
        skip_misc();
        auto root = element();
        return root;
    }
};

}  // namespace lbm

template <typename T>
class CConfiguration {
public:
    CVector<3, int> domain_size;
    CVector<3, int> subdomain_num;
    CVector<3, T> domain_length;
    CVector<3, T> gravitation;
    T viscosity = 0;
    CVector<4, T> drivenCavityVelocity;
    size_t computation_kernel_count = 128;
    int device_nr = 0;
    bool do_visualization = false;
    T timestep = -1;
    int loops = -1;
    bool do_validate = false;
    std::list<int> lbm_opencl_number_of_registers_list;
    std::list<int> lbm_opencl_number_of_threads_list;
    bool debug_mode = false;
    bool benchmark_output = false;

    // lattice-unit extension (optional <lattice> block)
    lbm::CLatticeOverrides<T> lattice;
    bool periodic[3] = {false, false, false};
    double perturbation_amplitude = 0.0;
    unsigned long long perturbation_seed = 0;
    int exchange_mode = 0;  // 0: all 19 populations after alpha (reference), 1: the 5 an inner cell reads
    // x faces: 1 inside the step kernel, 0 the copy launch (same results), 2 (LBM_FUSED_X_AUTO) the
    // faster of the two for the storage: inside the kernel for f32/f64; for f16/bf16 only the seam
    // wavefronts of a periodic ring of one, the copy launch elsewhere
    int fused_x = LBM_FUSED_X_AUTO;
    // alpha's own work-group size (lbm_group_set_alpha_work_group; 0: <kernel-count>)
    int alpha_work_group = 0;

    CConfiguration() {}
    explicit CConfiguration(const std::string &file_name) { loadFile(file_name); }

    // src/CConfiguration.hpp:69-78: throws on a file that cannot be read or parsed
    void loadFile(const std::string &file_name) {
        std::ifstream in(file_name);
        if (!in) throw "Loading XML file failed";
        std::stringstream ss;
        ss << in.rdbuf();
        std::unique_ptr<lbm::XmlNode> root;
        try {
            root = lbm::XmlReader(ss.str()).parse();
        } catch (const std::exception &) {
            throw "Loading XML file failed";
        }
        if (!root || root->name != "lbm-configuration") throw "Loading XML file failed";
        interpret(*root);
    }

    void printMe() const {
        std::cout << "################" << std::endl << "# CONFIGURATION " << std::endl << "################" << std::endl;
        std::cout << "PHYSICS: " << std::endl;
        std::cout << "\t    VISCOSITY: " << viscosity << std::endl;
        std::cout << "\t  GRAVITATION: " << gravitation << std::endl;
        std::cout << "     CAVITY VEL: " << drivenCavityVelocity << std::endl;
        std::cout << "GRID: " << std::endl;
        std::cout << "\t  DOMAIN_SIZE: " << domain_size << std::endl;
        std::cout << "\tSUBDOMIAN_NUM: " << subdomain_num << std::endl;
        std::cout << "SIMULATION: " << std::endl;
        std::cout << "\t        LOOPS: " << loops << std::endl;
        std::cout << "\t     TIMESTEP: " << timestep << std::endl;
        std::cout << "\t          VTK: " << do_visualization << std::endl;
        std::cout << "\t     VALIDATE: " << do_validate << std::endl;
        std::cout << "DEVICE: " << std::endl;
        std::cout << "  KERNEL_COUNT: " << computation_kernel_count << std::endl;
        std::cout << "\t    DEVICE_NR: " << device_nr << std::endl;
    }

private:
    static const lbm::XmlNode &need(const lbm::XmlNode *n, const char *name) {
        const lbm::XmlNode *c = n ? n->child(name) : nullptr;
        if (!c) throw "Loading XML file failed";
        return *c;
    }
    static double num(const lbm::XmlNode &n) { return atof(n.text.c_str()); }
    template <int N, typename V>
    static void vec(const lbm::XmlNode &n, CVector<N, V> &out, const char *const *names) {
        for (int i = 0; i < N; ++i) out[i] = (V)num(need(&n, names[i]));
    }

    void interpret(const lbm::XmlNode &root) {
        static const char *xyz[] = {"x", "y", "z"};
        static const char *xyzw[] = {"x", "y", "z", "w"};
        const lbm::XmlNode &dev = need(&root, "device");  // src/CConfiguration.hpp:146-150
        computation_kernel_count = (size_t)atoi(need(&dev, "kernel-count").text.c_str());
        device_nr = atoi(need(&dev, "device-number").text.c_str());
        const lbm::XmlNode &grid = need(&root, "grid");  // :125-138
        vec<3, int>(need(&grid, "domain-size"), domain_size, xyz);
        vec<3, int>(need(&grid, "subdomain-num"), subdomain_num, xyz);
        vec<3, T>(need(&grid, "domian-length"), domain_length, xyz);
        const lbm::XmlNode &phys = need(&root, "physics");  // :109-124
        viscosity = (T)num(need(&phys, "viscosity"));
        vec<3, T>(need(&phys, "gravitation"), gravitation, xyz);
        vec<4, T>(need(&phys, "cavity-velocity"), drivenCavityVelocity, xyzw);
        const lbm::XmlNode &sim = need(&root, "simulation");  // :139-145
        loops = atoi(need(&sim, "loops").text.c_str());
        timestep = (T)num(need(&sim, "timestep"));
        do_visualization = atoi(need(&need(&sim, "visualization"), "VTK").text.c_str()) != 0;
        do_validate = atoi(need(&sim, "validate").text.c_str()) != 0;
        if (const lbm::XmlNode *b = sim.child("benchmark")) benchmark_output = atoi(b->text.c_str()) != 0;
        if (const lbm::XmlNode *lat = root.child("lattice")) {
            lattice.enabled = true;
            if (auto *c = lat->child("tau")) lattice.tau = num(*c);
            if (auto *c = lat->child("lid-velocity")) lattice.lid_velocity = num(*c);
            if (auto *c = lat->child("body-force"))
                for (int i = 0; i < 3; ++i)
                    if (auto *d = c->child(xyz[i])) lattice.body_force[i] = num(*d);
            if (auto *c = lat->child("smagorinsky-cs")) lattice.smagorinsky_cs = num(*c);
            if (auto *c = lat->child("forcing")) lattice.forcing = c->text == "body" ? LBM_FORCING_BODY : LBM_FORCING_REFERENCE;
            if (auto *c = lat->child("kernel-variant")) lattice.kernel_variant = atoi(c->text.c_str());
            if (auto *c = lat->child("exchange-mode")) exchange_mode = atoi(c->text.c_str());
            if (auto *c = lat->child("fused-x")) fused_x = atoi(c->text.c_str());
            if (auto *c = lat->child("alpha-work-group")) alpha_work_group = atoi(c->text.c_str());
            if (auto *c = lat->child("storage")) {
                const std::string &v = c->text;
                lattice.storage = v == "f64" ? LBM_STORAGE_F64 : v == "f16" ? LBM_STORAGE_F16_SHIFTED
                                : v == "bf16" ? LBM_STORAGE_BF16_SHIFTED : LBM_STORAGE_F32;
            }
            if (auto *c = lat->child("periodic"))
                for (int i = 0; i < 3; ++i)
                    if (auto *d = c->child(xyz[i])) periodic[i] = atoi(d->text.c_str()) != 0;
            if (auto *c = lat->child("perturbation")) {
                if (auto *d = c->child("amplitude")) perturbation_amplitude = num(*d);
                if (auto *d = c->child("seed")) perturbation_seed = strtoull(d->text.c_str(), nullptr, 10);
            }
        }
    }
};

// src/Singleton.hpp:6-25 / src/common.h:59
template <typename C>
class Singleton {
public:
    static C *Instance() {
        static C instance;
        return &instance;
    }
};
