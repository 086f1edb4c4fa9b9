// builders_check.cpp -- test program: builds scene trees with the host layer's shape builders
// (sjfh::Box/Plane/.../assemble) and prints in() of every root for points read from stdin.
//
//   builders_check test                                  the Composite of the reference's
//                                                        tests/test.geom (offset 0.2, sum_list 10)
//   builders_check bowtie left rght mid top bot length   the two roots of the Au/SiO2 bowtie
//                                                        junction, in the reference's root order
//                                                        (the last defined Composite first)
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "sjf_host.hpp"

using namespace sjfh;

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string which = argv[1];
    std::vector<std::unique_ptr<Node>> roots;
    if (which == "test") {
        const double offset = 0.2, sum_list = 10;
        roots.push_back(assemble(Combine::Union, {
            Union({Box({.4, .4, .4}, {0.5, .5, .5}), Sphere({.4, .4, .6}, .1)}),
            Intersect({
                Difference({Box({.6, .4, .4}, {.6 + offset, .6, 0.4 + offset}),
                            Plane(Vec3{.6, .4, .6}, Vec3{8 / sum_list, .4, .6}, Vec3{.6, .6, .4})}),
                Cylinder({0.7, .5, .4}, .2, .1)}),
        }));
    } else if (which == "bowtie" && argc >= 8) {
        const double left = atof(argv[2]), rght = atof(argv[3]), mid = atof(argv[4]), top = atof(argv[5]),
                     bot = atof(argv[6]), length = atof(argv[7]);
        // SiO2 (defined second, first root)
        roots.push_back(assemble(Combine::Union, {
            Intersect({
                Intersect({
                    Union({Plane(Vec3{left - mid, 0, top}, Vec3{left, mid, bot}, Vec3{left, mid, top}),
                           Plane(Vec3{left - mid, length, top}, Vec3{left, mid, top}, Vec3{left, mid, bot})}),
                    Union({Plane(Vec3{rght + mid, 0, top}, Vec3{rght, mid, top}, Vec3{rght, mid, bot}),
                           Plane(Vec3{rght + mid, length, top}, Vec3{rght, mid, bot}, Vec3{rght, mid, top})}),
                }),
                Box({0, 0, top}, {length, length, bot}),
            }),
            Box({0, 0, bot}, {length, length, length}),
        }));
        // Au (defined first, second root)
        roots.push_back(assemble(Combine::Union, {
            Intersect({
                Union({
                    Intersect({Plane(Vec3{left - mid, 0, top}, Vec3{left, mid, top}, Vec3{left, mid, bot}),
                               Plane(Vec3{left - mid, length, top}, Vec3{left, mid, bot}, Vec3{left, mid, top})}),
                    Intersect({Plane(Vec3{rght + mid, 0, top}, Vec3{rght, mid, bot}, Vec3{rght, mid, top}),
                               Plane(Vec3{rght + mid, length, top}, Vec3{rght, mid, top}, Vec3{rght, mid, bot})}),
                }),
                Box({0, 0, top}, {length, length, bot}),
            }),
        }));
    } else {
        return 2;
    }
    double x, y, z;
    while (std::scanf("%lf %lf %lf", &x, &y, &z) == 3) {
        for (size_t r = 0; r < roots.size(); ++r) std::printf(r ? " %d" : "%d", in(*roots[r], Vec3{x, y, z}));
        std::printf("\n");
    }
    return 0;
}
