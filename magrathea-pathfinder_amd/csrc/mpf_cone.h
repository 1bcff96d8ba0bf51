// Cone geometry and photon launch of the reference programs (host):
// Cone::inside (cone.h:359-373), the spiral points of HyperSphere::uniform
// (abstracthypersphere.h:365-418), GenerateFullskyCones / GenerateNarrowCones
// (miscellaneous.h:200-402), the closest-cone selection of getTargets
// (miscellaneous.h:523-600), Hmaps::getPixels_per_cone2 (hmaps.h:261-322) and
// Integrator::launch (integrator.h:124-382).  Same operations in the same
// order, so cones, selections and launched photons are bit-identical to the
// reference's.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mpf {
namespace cone {

// vertex (3), base centre (3), half-angle (rad): the three members of
// Cone<std::array<double,3>> (cone.h data<0>, data<1>, data<2>)
struct Cone {
    double vertex[3];
    double base[3];
    double angle;
};

// Cone::inside (cone.h:359-373)
bool inside(const Cone& c, const double p[3]);

// The closest-cone test shared by getTargets, getPixels_per_cone2 and the
// launch into a series of cones: p is inside cones[self] and no other cone
// (in front of p) has its axis closer to p.  A cone is "other" when its index
// differs (by_value false: getPixels_per_cone2, launch) or when any of its
// members differs (by_value true: getTargets compares cones by value).
bool owns(const Cone* cones, uint32_t ncones, uint32_t self, const double p[3], bool by_value);

// Points of HyperSphere::uniform<2> on a sphere of the given radius centred
// at the origin (a spiral with the polar axis along y), xyz: n x 3
void spiral(uint64_t n, double radius, double* xyz);
// smallest distance between two of the first n points (the first member of
// the pair uniform returns; `init` when n < 2)
double min_distance(const double* xyz, uint64_t n, double init);

// GenerateFullskyCones (miscellaneous.h:200-218): cone i has its vertex at
// the origin, its base at spiral point i, and half-angle
// 1.8 asin(dmin / (2 radius))
void fullsky(uint32_t ncones, double radius, std::vector<Cone>& cones);

// GenerateNarrowCones (miscellaneous.h:257-402) for angles in degrees read
// from info_narrow_cone.txt: cones in the simulation frame, the same cones
// with their base rotated into the octree frame, and the rotation rotm1
// (row-major) that takes simulation coordinates into the octree frame.
// false with a message when no spiral size puts exactly ncones points in the
// region (the reference loops for ever there).
bool narrow(uint32_t ncones, double radius, double phi_deg, double theta_deg, double thetay_deg, double thetaz_deg,
            std::vector<Cone>& cones, std::vector<Cone>& cones_rot, double rotm1[9], uint64_t& spiral_size,
            std::string& err);

// Input::parse (input.h) of info_narrow_cone.txt and the conversions of
// miscellaneous.h:299-305 (phi through std::stoul)
bool narrow_info(const std::string& path, double& phi_deg, double& theta_deg, double& thetay_deg,
                 double& thetaz_deg, std::string& err);

// Utility::invMatrix3d (utility.h:704-719), row-major
void inv3(const double a[9], double r[9]);

// Photon launch of Integrator::launch; a photon is 9 doubles
// (a, t, x, y, z, dt/dl, dx/dl, dy/dl, dz/dl).
// From a point towards angles phi, theta (integrator.h:304-329)
void launch_angles(const double begin[3], double phi, double theta, double out9[9]);
// From a point towards another (integrator.h:264-290)
void launch_towards(const double begin[3], const double end[3], double out9[9]);
// rays.cpp:233-239 for `ray_targets = random`: with one std::mt19937 seeded
// with `seed` (seed + rank in the reference), n photons from the centre of a
// sphere of radius `micro_radius` towards random points of its surface that
// cones[icone] owns (integrator.h:166-249 with HyperSphere::random,
// abstracthypersphere.h:323-345), each followed by one uniform draw times
// 2 pi (the bundle rotation of the ray)
void launch_cone_random(const Cone* cones, uint32_t ncones, uint32_t icone, double micro_radius, uint32_t seed,
                        uint64_t n, double* out9, double* rotation);
// Photons on a circle of half-angle `angle` around one photon, turned by
// `rotation`, the photon itself first when `center` (integrator.h:345-382):
// count + center photons
void launch_bundle(const double in9[9], uint32_t count, double angle, double rotation, bool center, double* out9);

}  // namespace cone
}  // namespace mpf
