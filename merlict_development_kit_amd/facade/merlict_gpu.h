// merlict_gpu.h -- merlict's batch photon propagation on an MI355X.
//
// The reference-side binding of libmlgpu.so: compiled into a merlict build
// (it includes merlict's own headers and links libmlgpu.so and the HIP
// runtime), it takes merlict's types and gives merlict's results.
//
//   merlict::propagate_photons_in_frame_with_config        (merlict/Photons.h:12-16)
//   merlict::propagate_photons_in_frame_with_config_multi_thread
//                                                           (merlict_multi_thread.h:12-16)
//
// are replaced by the functions of the same name and signature in namespace
// merlict::gpu.  Each call flattens the initialised Frame tree of `world`
// into an mlgpu scene, propagates the photons on the GPU with one
// Philox4x32-10 stream per photon (key prng->create_seed(), stream = the
// photon's index in the vector), and writes every photon's history back:
// the Intersection and Interaction of each bounce
// (RayForPropagation::push_back_intersection_and_interaction) and the ray of
// its last leg.  Photon::time_of_flight, final_intersection,
// final_intersection_incident_direction_wrt_frame and
// sensor::Sensors::assign_photons then read the photons as they read
// photons propagated by merlict::Propagator.
//
// Errors: a wavelength outside a function table the photon meets raises
// std::out_of_range, as Func1::evaluate does (function/Func1.cpp:85-95); the
// photons before it are propagated, the failing one keeps the bounces before
// the failing one, the photons after it are untouched.  A scene the GPU does
// not model (a Frame subclass other than merlict's primitives, a frame with
// more than 16 children after init_tree_based_on_mother_child_relations, a
// propagation restriction to a frame outside the tree) raises
// std::invalid_argument before any photon is touched; a GPU failure raises
// std::runtime_error.
#ifndef MERLICT_GPU_H_
#define MERLICT_GPU_H_

#include <cstdint>
#include <memory>
#include <vector>

#include "merlict/Frame.h"
#include "merlict/Photon.h"
#include "merlict/PropagationConfig.h"
#include "merlict/random/Generator.h"

namespace merlict {
namespace gpu {

// merlict::propagate_photons_in_frame_with_config (merlict/Photons.cpp:9-22)
// on the GPU.
void propagate_photons_in_frame_with_config(
    std::vector<Photon>* photons,
    const Frame* world,
    const PropagationConfig* settings,
    random::Generator* prng);

// merlict::propagate_photons_in_frame_with_config_multi_thread
// (merlict_multi_thread.cpp:28-55) on the GPU: the same call, since the GPU
// runs every photon on a stream of its own anyway.
void propagate_photons_in_frame_with_config_multi_thread(
    std::vector<Photon>* photons,
    const Frame* world,
    const PropagationConfig* settings,
    random::Generator* prng);

// A world flattened onto the GPU once, for callers that propagate many
// batches through the same (unchanged) Frame tree.
class Scene {
 public:
    explicit Scene(const Frame* world);
    ~Scene();
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;

    // as propagate_photons_in_frame_with_config with this scene's world
    void propagate(
        std::vector<Photon>* photons,
        const PropagationConfig* settings,
        random::Generator* prng) const;

    // as propagate, with the Philox key given instead of drawn from a
    // Generator (photon i of the vector on stream i)
    void propagate_with_seed(
        std::vector<Photon>* photons,
        const PropagationConfig* settings,
        uint64_t seed) const;

    // index of a frame in the flattened scene (frame 0 is world->root()),
    // -1 for a frame outside it
    int32_t frame_index(const Frame* frame) const;
    int32_t num_frames() const;

    // photons per GPU batch (device memory ~ (150 + 64 x
    // max_num_interactions_per_photon) bytes each); default 1,048,576
    void set_batch_size(int64_t n);

 private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace gpu
}  // namespace merlict

#endif  // MERLICT_GPU_H_
