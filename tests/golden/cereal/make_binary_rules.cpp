// Writes cereal_binary_rules.bin: values of the kinds the binary scene
// schema relies on, encoded by the cereal release the reference vendors
// (cereal 1.1.2, build.md:88; include/cereal in the reference checkout) with
// its BinaryOutputArchive, as Scene::saveToFile does for ".bin" files
// (Scene.cpp:144-146).  Mock types only; the values are listed in
// cereal_binary_rules.json.  Build and run (in a checkout next to the
// reference):
//   g++ -std=c++14 -I<reference>/include make_binary_rules.cpp -o /tmp/mk && /tmp/mk cereal_binary_rules.bin
#include <cstdint>
#include <fstream>
#include <string>
#include <vector>

#include <cereal/archives/binary.hpp>
#include <cereal/types/base_class.hpp>
#include <cereal/types/string.hpp>
#include <cereal/types/vector.hpp>

enum class Kind { A, B, C };  // an enum class with the default underlying type, as SamplerType

struct V3 {
    double x = 0, y = 0, z = 0;
    template <class Ar>
    void serialize(Ar& ar) { ar(CEREAL_NVP(x), CEREAL_NVP(y), CEREAL_NVP(z)); }
};

struct Prim {  // polymorphic, as Primitive
    virtual ~Prim() {}
    uint64_t id = 0, materialId = 0;
    bool invisible = false;
    template <class Ar>
    void serialize(Ar& ar) { ar(CEREAL_NVP(id), CEREAL_NVP(materialId), CEREAL_NVP(invisible)); }
};

struct Tri : Prim {  // a base class by name, then a C array of structs, as Triangle
    V3 vertices[3];
    template <class Ar>
    void serialize(Ar& ar) { ar(cereal::make_nvp("primitive", cereal::base_class<Prim>(this)), CEREAL_NVP(vertices)); }
};

struct Rec {
    Kind kind = Kind::C;
    bool flag = true;
    int64_t offset = -5;
    uint64_t count = 0x0102030405060708ull;
    double value = 0.25;
    std::string path = "cube.obj";
    std::vector<Tri> tris;
    std::vector<uint64_t> ids{3, 1, 2};
    std::vector<Tri> none;
    template <class Ar>
    void serialize(Ar& ar) {
        ar(CEREAL_NVP(kind), CEREAL_NVP(flag), CEREAL_NVP(offset), CEREAL_NVP(count), CEREAL_NVP(value),
           CEREAL_NVP(path), CEREAL_NVP(tris), CEREAL_NVP(ids), CEREAL_NVP(none));
    }
};

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    Rec r;
    for (int t = 0; t < 2; ++t) {
        Tri tri;
        tri.id = 10 + t;
        tri.materialId = 20 + t;
        tri.invisible = t == 1;
        for (int v = 0; v < 3; ++v) tri.vertices[v] = V3{double(t), double(v), 0.5};
        r.tris.push_back(tri);
    }
    std::ofstream f(argv[1], std::ios::binary);
    cereal::BinaryOutputArchive ar(f);
    ar(cereal::make_nvp("scene", r));
    return 0;
}
