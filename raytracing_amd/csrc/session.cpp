// The session (see session.hpp): the scene's device copy, the view's double sums, the launch planner, and the two
// integrators behind one interface.
#include "session.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>

#include "device_scene.hpp"
#include "kernels/kernel_api.hpp"
#include "wavefront.hpp"

namespace rt
{
namespace
{
constexpr uint64_t round_seed_step = uint64_t{1} << 32;

class hip_event
{
public:
    // blocking: a wait for the event sleeps until the GPU signals instead of spinning (RT_WAIT_SLEEP)
    explicit hip_event(bool blocking = false)
    {
        check(hipEventCreateWithFlags(&e_, blocking ? hipEventBlockingSync : hipEventDefault), "creating a HIP event");
    }
    ~hip_event() { (void)hipEventDestroy(e_); }
    hip_event(const hip_event&) = delete;
    hip_event& operator=(const hip_event&) = delete;
    void record(hipStream_t s) { check(hipEventRecord(e_, s), "recording a HIP event"); }
    void synchronize() { check(hipEventSynchronize(e_), "rendering on the GPU"); }
    hipError_t query() const { return hipEventQuery(e_); }
    hipEvent_t get() const { return e_; }
    static double seconds_between(const hip_event& a, const hip_event& b)
    {
        float ms = 0;
        check(hipEventElapsedTime(&ms, a.e_, b.e_), "timing the GPU");
        return ms / 1000.0;
    }

private:
    hipEvent_t e_ = nullptr;
};

class stream_owner
{
public:
    stream_owner() { check(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "creating a HIP stream"); }
    ~stream_owner() { (void)hipStreamDestroy(s_); }
    stream_owner(const stream_owner&) = delete;
    stream_owner& operator=(const stream_owner&) = delete;
    hipStream_t get() const { return s_; }

private:
    hipStream_t s_ = nullptr;
};

}

class shared_scene
{
public:
    shared_scene(const rt_scene& scene, int32_t precision, int32_t tree, int32_t nodes, bool quality_kernel)
        : tables(scene_tables::copy_of(scene))
    {
        view = tables.view();
        if (precision != RT_FLOAT32)
        {
            f64 = std::make_unique<device_scene<double>>(view, tree, nodes, quality_kernel);
        }
        else
        {
            f32 = std::make_unique<device_scene<float>>(view, tree, nodes, quality_kernel);
        }
    }
    template <typename Real>
    const device_scene<Real>& device() const
    {
        if constexpr (sizeof(Real) == sizeof(double))
        {
            return *f64;
        }
        else
        {
            return *f32;
        }
    }
    scene_tables tables;
    rt_scene view{};
    std::unique_ptr<device_scene<float>> f32;
    std::unique_ptr<device_scene<double>> f64;
};

std::shared_ptr<const shared_scene> upload_scene(const rt_scene& scene, int32_t precision, int32_t tree, int32_t nodes,
                                                 int32_t kernel)
{
    validate_scene(scene);
    if (precision != RT_FLOAT32 && precision != RT_FLOAT64 && precision != RT_FLOAT64_ALL)
    {
        throw std::invalid_argument("unknown precision");
    }
    if (kernel < RT_KERNEL_AUTO || kernel > RT_KERNEL_QUALITY)
    {
        throw std::invalid_argument("unknown kernel");
    }
    const bool needed = scene.mesh_count != 0 || scene.microfacet_count != 0 || scene.environment != nullptr ||
                        scene.combination != RT_COMBINE_BSDF_ONLY;
    const bool quality = needed || kernel == RT_KERNEL_QUALITY;
    // the reference's refusals of the kernels asked for (src/gpu/renderer.cu:2070-2087)
    if (kernel == RT_KERNEL_ALL_BOOKS && needed)
    {
        throw std::invalid_argument("the quality kernel and the kernel for all three books can't both render");
    }
    if (kernel == RT_KERNEL_QUALITY && precision == RT_FLOAT64)
    {
        throw std::invalid_argument("the quality kernel computes in float only");
    }
    // (the reference's quality kernel is float32 only, src/gpu/quality.cu:53-109, tests/routing_tests.cu:53-67;
    // RT_FLOAT64_ALL renders it in float64 for an oracle)
    if (precision == RT_FLOAT64 && quality)
    {
        throw std::invalid_argument("triangles, GGX, environment maps and light samples render in float32 only");
    }
    // the quality integrator draws its own light samples: the lights list's strategies are the books' (the reference's
    // quality kernel refuses them too, tests/routing_tests.cu:53-67)
    if (quality && scene.camera.sampling != RT_SAMPLING_MATERIAL)
    {
        throw std::invalid_argument("triangles, GGX, environment maps and light samples take -p material, nee or mis, "
                                    "not the lights list's uniform, lights or mixture");
    }
    return std::make_shared<const shared_scene>(scene, precision, tree, nodes, quality);
}

std::size_t device_bytes(const shared_scene& scene) { return scene.f32 ? scene.f32->bytes() : scene.f64->bytes(); }

const rt_scene& scene_of(const shared_scene& scene) { return scene.view; }

namespace
{
// The passes the megakernel's staging buffer holds: at most 1 GiB of staged samples and 1024 passes, or `cap` passes
// (a serving session's step) when that is fewer
template <typename Real>
std::size_t staged_passes(int width, int height, int cap)
{
    constexpr std::size_t staging_bytes_cap = std::size_t{1} << 30;
    const std::size_t per_pass = megakernel_staging_per_pass(width, height);
    std::size_t passes = std::clamp<std::size_t>(staging_bytes_cap / (per_pass * sizeof(dev::staged_sample<Real>)), 1, 1024);
    if (cap > 0)
    {
        passes = std::min(passes, static_cast<std::size_t>(cap));
    }
    return passes;
}

// The quality megakernel's stacks go partly to HBM on a tree deeper than this (session_impl::ensure_far_stack)
constexpr int far_stack_depth = 16;
}

std::size_t view_device_bytes(const shared_scene& scene, const rt_camera_desc& camera, int staging_passes)
{
    const std::size_t pixels = static_cast<std::size_t>(camera.image_width) * static_cast<std::size_t>(camera.image_height);
    const bool f32 = scene.f32 != nullptr;
    const std::size_t staged = f32 ? staged_passes<float>(camera.image_width, camera.image_height, staging_passes) *
                                         sizeof(dev::staged_sample<float>)
                                   : staged_passes<double>(camera.image_width, camera.image_height, staging_passes) *
                                         sizeof(dev::staged_sample<double>);
    std::size_t bytes = 3 * sizeof(double) * pixels                                        // the sums
                        + staged * megakernel_staging_per_pass(camera.image_width, camera.image_height) // staging
                        + 2 * sizeof(unsigned long long) + sizeof(uint32_t);               // counters, queue head
    if (f32 && scene.f32->quality() && scene.f32->view().stack_depth > far_stack_depth)
    {
        bytes += megakernel_far_stack_bytes(scene.f32->view().stack_depth);
    }
    return bytes;
}

namespace
{
template <typename Real>
class session_impl final : public session
{
public:
    session_impl(std::shared_ptr<const shared_scene> scene, const rt_camera_desc& camera,
                 const rt_session_options& options)
        : shared_(std::move(scene)), scene_(shared_->tables), options_(options), dscene_(shared_->device<Real>()),
          start_(options.wait == RT_WAIT_SLEEP), stop_(options.wait == RT_WAIT_SLEEP)
    {
        view_scene_ = shared_->view;
        if (options.kernel == RT_KERNEL_QUALITY && !dscene_.quality())
        {
            throw std::invalid_argument("the resident scene has no quality tables for the quality kernel");
        }
        if (options.kernel == RT_KERNEL_ALL_BOOKS && dscene_.quality())
        {
            throw std::invalid_argument("the quality kernel and the kernel for all three books can't both render");
        }
        mode_ = dscene_.quality()                                                         ? dev::mode_quality
                : options.kernel != RT_KERNEL_ALL_BOOKS && first_book_scene(view_scene_) ? dev::mode_first_book
                                                                                          : dev::mode_all_books;
        launch_seconds_ = options.launch_seconds > 0 ? options.launch_seconds : 0.05;
        if (!megakernel_waves_supported(options.megakernel_waves))
        {
            throw std::invalid_argument("the megakernel's waves per SIMD must be 0, 5, 6 or 7");
        }
        if (options.integrator == RT_INTEGRATOR_WAVEFRONT && !wavefront_supported<Real>())
        {
            throw std::invalid_argument("the wavefront integrator renders in float32 only");
        }
        current_ = {options.integrator == RT_INTEGRATOR_WAVEFRONT, options.megakernel_waves,
                    options.block_size > 0 ? options.block_size : 64, options.integrator == RT_INTEGRATOR_PLAIN};
        if (options.integrator == RT_INTEGRATOR_AUTO && wavefront_supported<Real>())
        {
            // the autotuner's candidates, in the order they render the first full passes
            // (one wave a block: the megakernel's waves are independent, and a bigger block only holds more LDS)
            for (const int waves : {0, 5, 6, 7})
            {
                candidates_.push_back({false, waves, 64});
            }
            candidates_.push_back({true, 0, 256});
            current_ = candidates_.front();
            // the first round: every candidate renders a few passes (a screen), enough that the launch's chunk queue
            // holds ~4 chunks for every wave the chip can hold (2048 threads a CU): with fewer, a budget of more waves
            // wins the screen on its shorter tail alone (one pass of C1 screened out the compiler's budget, which
            // renders C1 -p mis 9 % faster than the budget picked)
            int cus = 0, device = 0;
            if (hipGetDevice(&device) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1)
            {
                cus = 256;
            }
            const double pixels =
                std::max(1.0, static_cast<double>(camera.image_width) * static_cast<double>(camera.image_height));
            const int screen = static_cast<int>(std::clamp(std::ceil(4.0 * cus * 2048.0 / pixels), 1.0, 16.0));
            // the session's first launch pays one-time costs (the kernels' code loaded, the staging buffer's pages
            // first touched): it renders the first candidate's passes untimed, and the screen follows
            schedule_.push_back({0, screen, false});
            for (std::size_t k = 0; k < candidates_.size(); ++k)
            {
                schedule_.push_back({k, screen});
            }
            tune_seconds_.assign(candidates_.size(), 0.0);
        }
        rays_ = device_array<unsigned long long>(2);
        queue_ = device_array<uint32_t>(1);
        validate_camera(camera);
        if (camera.sampling != scene_.camera.sampling)
        {
            throw std::invalid_argument("a view must keep the scene's sampling strategy");
        }
        start_view(camera);
    }

    ~session_impl() override
    {
        if (pending_ != 0)
        {
            (void)hipEventSynchronize(done_.get()); // a step in flight ends before its buffers are freed
        }
    }

    void set_camera(const rt_camera_desc& camera) override
    {
        settle();
        validate_camera(camera);
        if (camera.sampling != scene_.camera.sampling)
        {
            throw std::invalid_argument("a new view must keep the scene's sampling strategy");
        }
        has_view_ = false;
        start_view(camera);
    }

    accumulation state() const override { return state_; }
    rt_camera_desc camera() const override { return viewed_; }
    bool has_view() const override { return has_view_; }

    rt_render_report render_to(int64_t target_passes, int32_t target_rows,
                               const std::function<void(int64_t done)>& on_pass = {}) override
    {
        settle();
        require_view();
        const int height = camera_.image_height;
        if (target_passes < 0 || target_rows < 0 || target_rows >= height ||
            target_passes < state_.passes || (target_passes == state_.passes && target_rows < state_.rows))
        {
            throw std::invalid_argument("the target is behind the image or past its rows");
        }
        rt_render_report report{};
        while (state_.passes < target_passes || (state_.passes == target_passes && state_.rows < target_rows))
        {
            dev::launch_work work{};
            const int64_t pass = state_.passes;
            const int64_t round_size = static_cast<int64_t>(camera_.sqrt_spp) * camera_.sqrt_spp;
            const int64_t round = pass / round_size;
            work.first_pass = static_cast<int32_t>(pass % round_size);
            if (state_.rows > 0 || pass == target_passes)
            {
                // a band of rows of the pass in progress
                const int end_row = pass < target_passes ? height : target_rows;
                work.pass_count = 1;
                work.first_row = state_.rows;
                work.row_count = end_row - state_.rows;
            }
            else
            {
                const int64_t left_in_round = round_size - work.first_pass;
                int64_t count = seconds_per_pass_ > 0 ? static_cast<int64_t>(launch_seconds_ / seconds_per_pass_) : 1;
                // and no more passes than the megakernel's staging holds (one kernel a launch)
                const int64_t staged = current_.wavefront || current_.plain ? count : max_staged_passes();
                count = std::max<int64_t>(1, std::min({count, staged, left_in_round, target_passes - pass}));
                work.pass_count = tuning() ? static_cast<int32_t>(std::min<int64_t>(schedule_[tuned_].passes, std::min(
                                                 left_in_round, target_passes - pass)))
                                           : static_cast<int32_t>(count);
                work.first_row = 0;
                work.row_count = height;
            }
            const bool whole = work.first_row == 0 && work.row_count == height;
            const bool probe = whole && tuning();
            const variant v = probe ? candidates_[schedule_[tuned_].candidate] : current_;
            const double seconds = launch(options_.seed + static_cast<uint64_t>(round) * round_seed_step, work, v);
            report.launches += 1;
            report.bounces += last_bounces_;
            report.gpu_seconds += seconds;
            report.longest_launch_seconds = std::max(report.longest_launch_seconds, seconds);
            const int64_t passes_before = state_.passes;
            if (whole)
            {
                seconds_per_pass_ = seconds / work.pass_count;
                state_.passes += work.pass_count;
                if (probe)
                {
                    if (schedule_[tuned_].timed)
                    {
                        tune_seconds_[schedule_[tuned_].candidate] += seconds / work.pass_count;
                    }
                    if (++tuned_ == schedule_.size())
                    {
                        next_tuning_round();
                    }
                }
            }
            else if (work.first_row + work.row_count == height)
            {
                state_.passes += 1;
                state_.rows = 0;
            }
            else
            {
                state_.rows = work.first_row + work.row_count;
            }
            if (on_pass && state_.passes > passes_before)
            {
                on_pass(state_.passes); // (launch() has waited for the launch)
            }
        }
        report.passes = state_.passes;
        report.rows = state_.rows;
        return report;
    }

    rt_render_report render_to_static(int64_t target_passes) override
    {
        require_view();
        settle();
        if (target_passes < state_.passes || state_.rows != 0)
        {
            throw std::invalid_argument("the target is behind the image, or a pass is in progress");
        }
        std::size_t bytes = 0;
        void* const scratch = static_scratch(&bytes);
        if (scratch == nullptr)
        {
            throw std::runtime_error("the static scratch's device address is not available");
        }
        // the queue's head in the first 256 bytes, then the staged samples: one pass of the image must fit
        constexpr std::size_t head_bytes = 256;
        const std::size_t staged = (bytes - head_bytes) / sizeof(dev::staged_sample<Real>);
        const std::size_t per_pass = megakernel_staging_per_pass(camera_.image_width, camera_.image_height);
        if (per_pass > staged)
        {
            throw std::invalid_argument("an image of more than " + std::to_string(staged) +
                                        " pixels doesn't fit the static scratch (64 MiB)");
        }
        auto* const head = static_cast<uint32_t*>(scratch);
        void* const staging = static_cast<unsigned char*>(scratch) + head_bytes;
        // the megakernel of the session's register budget; the wavefront and the autotuner's candidates need memory
        // of their own, and every variant gives the same bytes
        const int waves = current_.wavefront ? 0 : current_.waves;
        rt_render_report report{};
        const int64_t round_size = static_cast<int64_t>(camera_.sqrt_spp) * camera_.sqrt_spp;
        while (state_.passes < target_passes)
        {
            dev::launch_work work{};
            const int64_t pass = state_.passes;
            work.first_pass = static_cast<int32_t>(pass % round_size);
            // (launch_megakernel runs a launch of more passes than the scratch stages as several kernels)
            work.pass_count = static_cast<int32_t>(std::min(target_passes - pass, round_size - work.first_pass));
            work.first_row = 0;
            work.row_count = camera_.image_height;
            const uint64_t seed = options_.seed + static_cast<uint64_t>(pass / round_size) * round_seed_step;
            start_.record(stream());
            check(launch_megakernel<Real>(mode_, cam_, dscene_.view(), seed, work, sums_.data(), staging, staged, head,
                                          rays_.data(), 64, stream(), nullptr, waves),
                  "launching the kernel");
            stop_.record(stream());
            stop_.synchronize();
            const double seconds = hip_event::seconds_between(start_, stop_);
            kernel_seconds_ += seconds;
            report.launches += 1;
            report.gpu_seconds += seconds;
            report.longest_launch_seconds = std::max(report.longest_launch_seconds, seconds);
            state_.passes += work.pass_count;
        }
        report.passes = state_.passes;
        report.rows = state_.rows;
        return report;
    }

    rt_render_report render_for(double seconds, int64_t pass_limit) override
    {
        require_view();
        if (!(seconds >= 0 && seconds <= 3600) || pass_limit < 0)
        {
            throw std::invalid_argument("a time from 0 to 3600 seconds and a pass limit of 0 or more");
        }
        rt_render_report report{};
        // a full launch: the session's planned length, or the whole time when that is less
        const double full = std::min(seconds, launch_seconds_);
        while (state_.passes < pass_limit && (report.gpu_seconds < seconds || state_.passes == 0))
        {
            // after a first launch, end early rather than make one shorter than a quarter of a full one
            // (src/gpu/renderer.h:606-612)
            if (report.launches > 0 && state_.passes > 0 && seconds - report.gpu_seconds < 0.25 * full)
            {
                break;
            }
            const int64_t remaining = pass_limit - state_.passes;
            int64_t step = seconds_per_pass_ > 0
                               ? static_cast<int64_t>(std::max(1.0, (seconds - report.gpu_seconds) / seconds_per_pass_))
                               : 1;
            step = std::min(step, remaining);
            const rt_render_report r = render_to(state_.passes + step, 0);
            report.launches += r.launches;
            report.bounces += r.bounces;
            report.gpu_seconds += r.gpu_seconds;
            report.longest_launch_seconds = std::max(report.longest_launch_seconds, r.longest_launch_seconds);
        }
        report.passes = state_.passes;
        report.rows = state_.rows;
        return report;
    }

    void pixels(double* out, std::size_t count) override
    {
        require_view();
        if (state_.passes == 0 || state_.rows != 0)
        {
            throw std::logic_error("the image has no complete pass, or a pass in progress");
        }
        sums(out, count);
        const double inverse = 1.0 / static_cast<double>(state_.passes);
        for (std::size_t i = 0; i < values(); ++i)
        {
            out[i] *= inverse;
        }
    }

    void sums(double* out, std::size_t count) override
    {
        settle();
        require_view();
        if (count < values())
        {
            throw std::invalid_argument("the output buffer is smaller than width x height x 3");
        }
        check(hipStreamSynchronize(stream()), "waiting for the GPU");
        check(hipMemcpy(out, sums_.data(), values() * sizeof(double), hipMemcpyDeviceToHost), "copying the image");
    }

    void display(uint8_t* out, std::size_t count) override
    {
        settle();
        require_view();
        if (state_.passes == 0)
        {
            throw std::logic_error("the image has no complete pass");
        }
        const std::size_t pixels = values() / 3;
        if (count < 4 * pixels)
        {
            throw std::invalid_argument("the output buffer is smaller than width x height x 4");
        }
        if (rgba_.size() != 4 * pixels)
        {
            rgba_ = device_array<uint8_t>(4 * pixels);
        }
        check(launch_display(sums_.data(), static_cast<int64_t>(pixels), 1.0 / static_cast<double>(state_.passes),
                             rgba_.data(), stream()),
              "launching the display kernel");
        check(hipStreamSynchronize(stream()), "computing the display");
        check(hipMemcpy(out, rgba_.data(), 4 * pixels, hipMemcpyDeviceToHost), "copying the display");
    }

    void display_toned(int32_t tone_map, double exposure, uint8_t* out, std::size_t count) override
    {
        if (tone_map < RT_TONE_BOOKS || tone_map > RT_TONE_NEUTRAL)
        {
            throw std::invalid_argument("unknown tone map");
        }
        if (!std::isfinite(exposure) || std::fabs(exposure) > RT_MAX_EXPOSURE)
        {
            throw std::invalid_argument("the exposure must be a finite number of stops within +-24");
        }
        if (tone_map == RT_TONE_BOOKS)
        {
            if (exposure != 0)
            {
                throw std::invalid_argument("the books' bytes take no exposure");
            }
            display(out, count);
            return;
        }
        if constexpr (sizeof(Real) != sizeof(float))
        {
            throw std::invalid_argument("tone maps other than the books' compute in float32 sessions only");
        }
        require_view();
        if (state_.passes == 0)
        {
            throw std::logic_error("the image has no complete pass");
        }
        const std::size_t pixels = values() / 3;
        if (count < 4 * pixels)
        {
            throw std::invalid_argument("the output buffer is smaller than width x height x 4");
        }
        if (rgba_.size() != 4 * pixels)
        {
            rgba_ = device_array<uint8_t>(4 * pixels);
        }
        check(launch_tone_map(sums_.data(), camera_.image_width, static_cast<int64_t>(pixels),
                              1.0 / static_cast<double>(state_.passes), tone_map,
                              static_cast<float>(std::exp2(exposure)), rgba_.data(), stream()),
              "launching the tone map");
        check(hipStreamSynchronize(stream()), "computing the display");
        check(hipMemcpy(out, rgba_.data(), 4 * pixels, hipMemcpyDeviceToHost), "copying the display");
    }

    void counters(uint64_t* rays, uint64_t* shadow_rays, double* kernel_seconds) const override
    {
        require_view();
        unsigned long long n[2] = {};
        check(hipMemcpy(n, rays_.data(), sizeof n, hipMemcpyDeviceToHost), "reading the ray counters");
        if (rays)
        {
            *rays = n[0];
        }
        if (shadow_rays)
        {
            *shadow_rays = n[1];
        }
        if (kernel_seconds)
        {
            *kernel_seconds = kernel_seconds_;
        }
    }

    rt_kernel_info kernel() const override
    {
        rt_kernel_info info{};
        kernel_attributes a{};
        const int depth = dscene_.view().stack_depth;
        if (current_.wavefront)
        {
            a = wf_attributes(depth, 0);
            const kernel_attributes s = wf_attributes(depth, mode_ == dev::mode_quality ? 2 : 1);
            a.registers = std::max(a.registers, s.registers);
            a.local_bytes = std::max(a.local_bytes, s.local_bytes);
            info.threads_per_block = dev::wf_block;
            std::snprintf(info.name, sizeof info.name, "wavefront");
        }
        else if (current_.plain)
        {
            a = plain_megakernel_attributes<Real>(mode_, depth, dev::tree_walk(dscene_.view()));
            const char* name =
                mode_ == dev::mode_quality ? "quality" : (mode_ == dev::mode_all_books ? "all books" : "first book");
            info.threads_per_block = plain_block_x * plain_block_y;
            std::snprintf(info.name, sizeof info.name, "plain %s%s %dx%d", name, sizeof(Real) == 8 ? " fp64" : "",
                          plain_block_x, plain_block_y);
        }
        else
        {
            a = megakernel_attributes<Real>(mode_, current_.block, depth, current_.waves,
                                            dev::tree_walk(dscene_.view()), dscene_.view().noises != nullptr);
            const char* name =
                mode_ == dev::mode_quality ? "quality" : (mode_ == dev::mode_all_books ? "all books" : "first book");
            info.threads_per_block = current_.block;
            char budget[16] = "";
            if (current_.waves > 0)
            {
                std::snprintf(budget, sizeof budget, " w%d", current_.waves);
            }
            std::snprintf(info.name, sizeof info.name, "%s%s%s b%d", name, sizeof(Real) == 8 ? " fp64" : "", budget,
                          current_.block);
        }
        info.registers = a.registers;
        info.local_bytes = a.local_bytes;
        info.lds_bytes = a.lds_bytes;
        info.blocks_per_cu = a.blocks_per_cu;
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) == hipSuccess)
        {
            info.compute_units = cus;
        }
        return info;
    }

    void closest_hit(const float* origins, const float* directions, const float* times, const int32_t* from,
                     const uint8_t* inward, int64_t n, int32_t* index, float* t, uint8_t* entering) override
    {
        if constexpr (sizeof(Real) != sizeof(float))
        {
            throw std::invalid_argument("the closest-hit probe runs in float32 sessions only");
        }
        else
        {
            if (n <= 0)
            {
                return;
            }
            const std::size_t m = static_cast<std::size_t>(n);
            device_array<float> o(3 * m), d(3 * m), tm(m), tt(m);
            device_array<int32_t> f(m), idx(m);
            device_array<uint8_t> in(m), ent(m);
            check(hipMemcpy(o.data(), origins, o.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(d.data(), directions, d.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(tm.data(), times, tm.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(f.data(), from, f.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(in.data(), inward, in.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(launch_closest_hit<float>(dscene_.view(), o.data(), d.data(), tm.data(), f.data(), in.data(), n,
                                            idx.data(), tt.data(), ent.data(), stream()),
                  "launching the closest-hit probe");
            check(hipStreamSynchronize(stream()), "tracing the probe rays");
            check(hipMemcpy(index, idx.data(), idx.bytes(), hipMemcpyDeviceToHost), "copying hits");
            check(hipMemcpy(t, tt.data(), tt.bytes(), hipMemcpyDeviceToHost), "copying hits");
            check(hipMemcpy(entering, ent.data(), ent.bytes(), hipMemcpyDeviceToHost), "copying hits");
        }
    }

    void occluded(const float* origins, const float* directions, const float* times, const int32_t* from,
                  const uint8_t* inward, const float* t_max, const int32_t* skip, int64_t n, uint8_t* blocked,
                  uint8_t* walked) override
    {
        if constexpr (sizeof(Real) != sizeof(float))
        {
            throw std::invalid_argument("the shadow-ray probe runs in float32 sessions only");
        }
        else
        {
            if (n <= 0)
            {
                return;
            }
            const std::size_t m = static_cast<std::size_t>(n);
            device_array<float> o(3 * m), d(3 * m), tm(m), tx(m);
            device_array<int32_t> f(m), sk(m);
            device_array<uint8_t> in(m), b(m), w(m);
            check(hipMemcpy(o.data(), origins, o.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(d.data(), directions, d.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(tm.data(), times, tm.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(f.data(), from, f.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(in.data(), inward, in.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(tx.data(), t_max, tx.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(sk.data(), skip, sk.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(launch_occluded(dscene_.view(), o.data(), d.data(), tm.data(), f.data(), in.data(), tx.data(),
                                  sk.data(), n, b.data(), w.data(), stream()),
                  "launching the shadow-ray probe");
            check(hipStreamSynchronize(stream()), "tracing the probe rays");
            if (blocked)
            {
                check(hipMemcpy(blocked, b.data(), b.bytes(), hipMemcpyDeviceToHost), "copying answers");
            }
            if (walked)
            {
                check(hipMemcpy(walked, w.data(), w.bytes(), hipMemcpyDeviceToHost), "copying answers");
            }
        }
    }

    void media_events(const float* origins, const float* directions, const float* times, const float* t_surface,
                      int64_t n, uint64_t seed, int32_t* medium, float* t, float* transmitted) override
    {
        if constexpr (sizeof(Real) != sizeof(float))
        {
            throw std::invalid_argument("the media probe runs in float32 sessions only");
        }
        else
        {
            if (n <= 0)
            {
                return;
            }
            const std::size_t m = static_cast<std::size_t>(n);
            device_array<float> o(3 * m), d(3 * m), tm(m), ts(m), tt(m), tr(m);
            device_array<int32_t> md(m);
            check(hipMemcpy(o.data(), origins, o.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(d.data(), directions, d.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(tm.data(), times, tm.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(ts.data(), t_surface, ts.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(launch_media_probe(dscene_.view(), o.data(), d.data(), tm.data(), ts.data(), n, seed, md.data(),
                                     tt.data(), tr.data(), stream()),
                  "launching the media probe");
            check(hipStreamSynchronize(stream()), "tracing the media");
            check(hipMemcpy(medium, md.data(), md.bytes(), hipMemcpyDeviceToHost), "copying events");
            check(hipMemcpy(t, tt.data(), tt.bytes(), hipMemcpyDeviceToHost), "copying events");
            check(hipMemcpy(transmitted, tr.data(), tr.bytes(), hipMemcpyDeviceToHost), "copying events");
        }
    }

    void light_samples(const float* points, int64_t n, uint64_t seed, float* wi, float* distance, float* radiance,
                       float* density, int32_t* skip, float* hit_density) override
    {
        if constexpr (sizeof(Real) != sizeof(float))
        {
            throw std::invalid_argument("the light-sample probe runs in float32 sessions only");
        }
        else
        {
            if (!dscene_.quality())
            {
                throw std::invalid_argument("the light-sample probe needs a scene with light samples (-p nee or mis, "
                                            "an environment map or GGX)");
            }
            if (!(dscene_.view().quality.light_pmf > 0.0f))
            {
                throw std::invalid_argument("the scene's light set is empty");
            }
            if (n <= 0)
            {
                return;
            }
            const std::size_t m = static_cast<std::size_t>(n);
            device_array<float> p(3 * m), w(3 * m), dist(m), rad(3 * m), den(m), hit(m);
            device_array<int32_t> sk(m);
            check(hipMemcpy(p.data(), points, p.bytes(), hipMemcpyHostToDevice), "copying points");
            check(launch_light_probe(dscene_.view(), p.data(), n, seed, w.data(), dist.data(), rad.data(), den.data(),
                                     sk.data(), hit.data(), stream()),
                  "launching the light-sample probe");
            check(hipStreamSynchronize(stream()), "drawing light samples");
            check(hipMemcpy(wi, w.data(), w.bytes(), hipMemcpyDeviceToHost), "copying samples");
            check(hipMemcpy(distance, dist.data(), dist.bytes(), hipMemcpyDeviceToHost), "copying samples");
            check(hipMemcpy(radiance, rad.data(), rad.bytes(), hipMemcpyDeviceToHost), "copying samples");
            check(hipMemcpy(density, den.data(), den.bytes(), hipMemcpyDeviceToHost), "copying samples");
            check(hipMemcpy(skip, sk.data(), sk.bytes(), hipMemcpyDeviceToHost), "copying samples");
            check(hipMemcpy(hit_density, hit.data(), hit.bytes(), hipMemcpyDeviceToHost), "copying samples");
        }
    }

    void reference_hits(const float* origins, const float* directions, const float* times, const int32_t* from,
                        const uint8_t* inward, int64_t n, int32_t* index, float* t, uint8_t* entering) override
    {
        if constexpr (sizeof(Real) != sizeof(float))
        {
            throw std::invalid_argument("the reference-walk probe runs in float32 sessions only");
        }
        else
        {
            if (n <= 0)
            {
                return;
            }
            if (dscene_.view().triangle_base != dscene_.view().media_base)
            {
                throw std::invalid_argument("the reference-walk probe walks trees of spheres and quads only");
            }
            const auto& nodes = dscene_.tree().nodes;
            static_assert(sizeof(ref_node<float>) == 32, "the probe reads the reference's 32-byte node");
            if (binary_node_bytes() != sizeof(ref_node<float>) || nodes.size() > static_cast<std::size_t>(INT32_MAX))
            {
                throw std::runtime_error("the binary tree's layout differs from the probe's");
            }
            if (binary_nodes_.size() != nodes.size())
            {
                binary_nodes_ = device_array<ref_node<float>>(nodes);
            }
            const std::size_t m = static_cast<std::size_t>(n);
            device_array<float> o(3 * m), d(3 * m), tm(m), tt(m);
            device_array<int32_t> f(m), idx(m), overflow(1);
            device_array<uint8_t> in(m), ent(m);
            check(hipMemcpy(o.data(), origins, o.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(d.data(), directions, d.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(tm.data(), times, tm.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(f.data(), from, f.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemcpy(in.data(), inward, in.bytes(), hipMemcpyHostToDevice), "copying rays");
            check(hipMemset(overflow.data(), 0, overflow.bytes()), "clearing the probe's counter");
            check(launch_reference_walk(dscene_.view(), binary_nodes_.data(), static_cast<int>(nodes.size()), o.data(),
                                        d.data(), tm.data(), f.data(), in.data(), n, idx.data(), tt.data(), ent.data(),
                                        overflow.data(), stream()),
                  "launching the reference-walk probe");
            check(hipStreamSynchronize(stream()), "tracing the probe rays");
            int32_t overflowed = 0;
            check(hipMemcpy(&overflowed, overflow.data(), sizeof overflowed, hipMemcpyDeviceToHost), "reading hits");
            if (overflowed != 0)
            {
                throw std::runtime_error("the reference walk's stack overflowed on some rays");
            }
            check(hipMemcpy(index, idx.data(), idx.bytes(), hipMemcpyDeviceToHost), "copying hits");
            check(hipMemcpy(t, tt.data(), tt.bytes(), hipMemcpyDeviceToHost), "copying hits");
            check(hipMemcpy(entering, ent.data(), ent.bytes(), hipMemcpyDeviceToHost), "copying hits");
        }
    }

    rt_traffic traffic(int64_t passes) override
    {
        settle();
        require_view();
        const int64_t round_size = static_cast<int64_t>(camera_.sqrt_spp) * camera_.sqrt_spp;
        if (passes < 1 || passes > round_size)
        {
            throw std::invalid_argument("traffic is counted over 1 to sqrt_spp^2 passes of the first round");
        }
        device_array<double> scratch(values());
        // [rays, shadow rays, the tally events, the phases' lane-cycles]
        device_array<unsigned long long> counts(2 + RT_TRAFFIC_EVENTS + RT_TRAFFIC_PHASES);
        check(hipMemsetAsync(scratch.data(), 0, scratch.bytes(), stream()), "clearing scratch");
        check(hipMemsetAsync(counts.data(), 0, counts.bytes(), stream()), "clearing counters");
        const dev::launch_work work{0, static_cast<int32_t>(passes), 0, camera_.image_height};
        ensure_staging();
        check(launch_megakernel<Real>(mode_, cam_, dscene_.view(), options_.seed, work, scratch.data(),
                                      staging_.data(), staging_samples_, queue_.data(), counts.data(), 128,
                                      stream(), counts.data() + 2),
              "launching the counting kernel");
        check(hipStreamSynchronize(stream()), "counting traffic");
        unsigned long long host[2 + RT_TRAFFIC_EVENTS + RT_TRAFFIC_PHASES] = {};
        check(hipMemcpy(host, counts.data(), sizeof host, hipMemcpyDeviceToHost), "reading counters");
        rt_traffic t{};
        t.paths = static_cast<uint64_t>(passes) * static_cast<uint64_t>(values() / 3);
        t.rays = host[0];
        t.shadow_rays = host[1];
        // (a packed tree's walks read 64-byte nodes)
        const double node_bytes = dscene_.view().packed_nodes != nullptr ? sizeof(dev::packed_node)
                                                                         : sizeof(dev::wide_node<Real>);
        const double sizes[RT_TRAFFIC_EVENTS] = {
            node_bytes, sizeof(dev::sphere_rec<Real>), sizeof(dev::moving_sphere_rec<Real>),
            sizeof(dev::quad_rec<Real>), sizeof(dev::medium_rec<Real>), sizeof(dev::quad_rec<Real>),
            4.0 + sizeof(dev::material_rec<Real>), sizeof(dev::sphere_frame<Real>), sizeof(dev::texture_rec<Real>),
            sizeof(dev::texel), 4.0 * sizeof(Real) + 3.0, sizeof(dev::light_rec<Real>),
            sizeof(dev::triangle_rec<Real>), sizeof(dev::emitter_rec), sizeof(dev::env_texel), sizeof(dev::alias_bin),
            sizeof(dev::instance_rec<Real>)};
        double bytes = 0;
        for (int e = 0; e < RT_TRAFFIC_EVENTS; ++e)
        {
            t.events[e] = host[2 + e];
            t.event_bytes[e] = sizes[e];
            bytes += sizes[e] * static_cast<double>(host[2 + e]);
        }
        for (int k = 0; k < RT_TRAFFIC_PHASES; ++k)
        {
            t.phase_lane_cycles[k] = host[2 + RT_TRAFFIC_EVENTS + k];
        }
        // the pixel's double sums, read and written once per launch of `passes` passes
        bytes += 48.0 * static_cast<double>(values() / 3);
        t.bytes_per_path = bytes / static_cast<double>(t.paths);
        return t;
    }

    std::size_t device_bytes() const override { return dscene_.bytes() + view_device_bytes(); }

    std::size_t view_device_bytes() const override
    {
        return sums_.bytes() + staging_.bytes() + far_stack_.bytes() + rays_.bytes() + queue_.bytes() + rgba_.bytes() +
               (wavefront_ ? wavefront_->bytes() : 0);
    }

    void serve_on(void* shared_stream, int staging_passes) override
    {
        if (staging_passes < 0)
        {
            throw std::invalid_argument("0 or more staged passes");
        }
        settle();
        check(hipStreamSynchronize(stream()), "waiting for the GPU");
        serve_stream_ = static_cast<hipStream_t>(shared_stream);
        if (staging_passes != staging_cap_)
        {
            staging_cap_ = staging_passes;
            staging_ = device_array<dev::staged_sample<Real>>(); // sized again at the next launch
            staging_samples_ = 0;
        }
    }

    void enqueue_to(int64_t target, int grid_ways) override
    {
        require_view();
        if (grid_ways < 1)
        {
            throw std::invalid_argument("a share of the chip of 1/n with n >= 1");
        }
        if (pending_ != 0)
        {
            throw std::logic_error("the session's last step hasn't been collected");
        }
        if (state_.rows != 0 || target <= state_.passes)
        {
            throw std::invalid_argument("a step adds whole passes to an image without a pass in progress");
        }
        if (current_.wavefront || current_.plain || tuning())
        {
            throw std::logic_error("steps render on the megakernel of a session that isn't tuning");
        }
        ensure_staging();
        ensure_far_stack();
        const int64_t round_size = static_cast<int64_t>(camera_.sqrt_spp) * camera_.sqrt_spp;
        for (int64_t pass = state_.passes; pass < target;)
        {
            dev::launch_work work{};
            work.first_pass = static_cast<int32_t>(pass % round_size);
            work.pass_count = static_cast<int32_t>(std::min(target - pass, round_size - work.first_pass));
            work.first_row = 0;
            work.row_count = camera_.image_height;
            const uint64_t seed = options_.seed + static_cast<uint64_t>(pass / round_size) * round_seed_step;
            check(launch_megakernel<Real>(mode_, cam_, dscene_.view(), seed, work, sums_.data(), staging_.data(),
                                          staging_samples_, queue_.data(), rays_.data(), current_.block, stream(),
                                          nullptr, current_.waves, far_stack_.data(), far_stack_.bytes(), grid_ways),
                  "launching the kernel");
            pass += work.pass_count;
        }
        done_.record(stream());
        pending_ = target;
    }

    bool poll() override
    {
        if (pending_ == 0)
        {
            return true;
        }
        const hipError_t e = done_.query();
        if (e == hipErrorNotReady)
        {
            return false;
        }
        check(e, "rendering on the GPU");
        state_.passes = pending_;
        pending_ = 0;
        return true;
    }

private:
    // serving: the stream the launches go to (null: the session's own), the staged passes a launch is capped at (0:
    // the default), and the passes of the step in flight (0: none)
    hipStream_t serve_stream_ = nullptr;
    int staging_cap_ = 0;
    int64_t pending_ = 0;
    hip_event done_;
    hipStream_t stream() const { return serve_stream_ ? serve_stream_ : stream_.get(); }
    // a step in flight waited for, and its passes added to the image's state
    void settle()
    {
        if (pending_ != 0)
        {
            done_.synchronize();
            poll();
        }
    }

    std::shared_ptr<const shared_scene> shared_;
    const scene_tables& scene_;
    rt_scene view_scene_{};
    rt_session_options options_;
    const device_scene<Real>& dscene_;
    stream_owner stream_;
    hip_event start_, stop_;
    int mode_ = dev::mode_all_books;
    // what renders a launch: the megakernel (its register budget and block size) or the wavefront
    struct variant
    {
        bool wavefront = false;
        int waves = 0;
        int block = 128;
        bool plain = false; // the reference's mapping: a thread per pixel, its passes in a loop
    };
    variant current_{};
    struct probe
    {
        std::size_t candidate;
        int passes;        // at most: a launch never crosses a round of strata or the render's target
        bool timed = true; // false: a launch that renders its passes but whose time is not compared
    };
    std::vector<variant> candidates_;  // the autotuner's
    std::vector<probe> schedule_;      // the first full launches: which candidate renders each, how many passes
    std::size_t tuned_ = 0;            // launches of the schedule rendered so far
    int tuning_round_ = 0;
    std::vector<double> tune_seconds_; // per candidate, the sum of its probes' seconds per pass this round
    bool tuning() const { return tuned_ < schedule_.size(); }

    // After the screen (one pass each), the three fastest candidates are timed again on launches of several passes
    // (a planned launch's worth, up to 256 passes: the launches they will render), in a forward then a backward sweep so that
    // a drift in the timings (clocks, caches warming) weighs on all of them alike; then the fastest renders the rest
    void next_tuning_round()
    {
        std::vector<std::size_t> order(candidates_.size());
        std::iota(order.begin(), order.end(), std::size_t{0});
        std::vector<double> per_pass(candidates_.size(), std::numeric_limits<double>::infinity());
        std::vector<int> probes(candidates_.size(), 0);
        for (const probe& p : schedule_)
        {
            probes[p.candidate] += p.timed ? 1 : 0;
        }
        for (std::size_t k = 0; k < candidates_.size(); ++k)
        {
            if (probes[k] > 0)
            {
                per_pass[k] = tune_seconds_[k] / probes[k];
            }
        }
        std::stable_sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) { return per_pass[a] < per_pass[b]; });
        if (tuning_round_ == 0 && candidates_.size() > 1)
        {
            tuning_round_ = 1;
            const std::size_t finalists = std::min<std::size_t>(3, candidates_.size());
            const int passes = static_cast<int>(
                std::clamp(launch_seconds_ / std::max(per_pass[order[0]], 1e-9), 1.0, 256.0));
            schedule_.clear();
            tuned_ = 0;
            for (std::size_t k = 0; k < finalists; ++k)
            {
                schedule_.push_back({order[k], passes});
            }
            for (std::size_t k = finalists; k-- > 0;)
            {
                schedule_.push_back({order[k], passes});
            }
            tune_seconds_.assign(candidates_.size(), 0.0);
            return;
        }
        current_ = candidates_[order[0]];
        seconds_per_pass_ = per_pass[order[0]];
    }
    std::unique_ptr<wavefront<Real>> wavefront_; // made at the first launch that needs it in a view
    double launch_seconds_ = 0.05;
    rt_camera_desc camera_{};
    dev::camera_rec<Real> cam_{};
    device_array<double> sums_;
    device_array<unsigned long long> rays_;
    // the megakernel's chunk queue head, and its staged samples: one per pixel per pass, for up to staging_passes_
    // passes a kernel
    device_array<uint32_t> queue_;
    device_array<dev::staged_sample<Real>> staging_;
    std::size_t staging_samples_ = 0;
    int staging_passes_ = 0;
    int64_t max_staged_passes() const
    {
        return static_cast<int64_t>(staged_passes<Real>(camera_.image_width, camera_.image_height, staging_cap_));
    }
    void ensure_staging()
    {
        if (staging_samples_ != 0)
        {
            return;
        }
        const std::size_t per_pass = megakernel_staging_per_pass(camera_.image_width, camera_.image_height);
        const std::size_t passes = staged_passes<Real>(camera_.image_width, camera_.image_height, staging_cap_);
        staging_ = device_array<dev::staged_sample<Real>>(per_pass * passes);
        staging_samples_ = per_pass * passes;
        staging_passes_ = static_cast<int>(passes);
    }
    // the quality megakernel's stacks' deeper levels on a deep tree (a large mesh's), allocated at its first launch
    device_array<uint8_t> far_stack_;
    void ensure_far_stack()
    {
        const int depth = dscene_.view().stack_depth;
        if (mode_ == dev::mode_quality && depth > far_stack_depth && far_stack_.size() == 0)
        {
            far_stack_ = device_array<uint8_t>(megakernel_far_stack_bytes(depth));
        }
    }
    device_array<uint8_t> rgba_;
    device_array<ref_node<float>> binary_nodes_; // the reference-walk probe's copy of the binary tree
    bool has_view_ = false;
    rt_camera_desc viewed_{}; // the last view set successfully
    accumulation state_{};
    double kernel_seconds_ = 0;
    double seconds_per_pass_ = 0;
    int64_t last_bounces_ = 0; // the wavefront's bounce steps in the last launch

    std::size_t values() const
    {
        return 3ull * static_cast<std::size_t>(camera_.image_width) * static_cast<std::size_t>(camera_.image_height);
    }
    void require_view() const
    {
        if (!has_view_)
        {
            throw std::logic_error("the session has no view: set a camera first");
        }
    }

    void start_view(const rt_camera_desc& camera)
    {
        camera_ = camera;
        cam_ = to_camera<Real>(camera);
        const std::size_t n = values();
        if (sums_.size() != n)
        {
            sums_ = device_array<double>();
            sums_ = device_array<double>(n);
        }
        check(hipMemsetAsync(sums_.data(), 0, sums_.bytes(), stream()), "clearing the image");
        check(hipMemsetAsync(rays_.data(), 0, rays_.bytes(), stream()), "clearing the counters");
        wavefront_.reset(); // its pool is the view's size
        if (staging_samples_ != megakernel_staging_per_pass(camera.image_width, camera.image_height) * staging_passes_)
        {
            staging_ = device_array<dev::staged_sample<Real>>(); // the megakernel's, sized again at its next launch
            staging_samples_ = 0;
        }
        check(hipStreamSynchronize(stream()), "starting the view");
        state_ = {};
        kernel_seconds_ = 0;
        seconds_per_pass_ = 0;
        has_view_ = true;
        viewed_ = camera;
    }

    // One launch of the view's work, timed on the GPU
    double launch(uint64_t seed, const dev::launch_work& work, const variant& v)
    {
        if (v.wavefront && !wavefront_)
        {
            wavefront_ = std::make_unique<wavefront<Real>>(dscene_.view(), cam_, mode_, options_, stream());
        }
        start_.record(stream());
        last_bounces_ = 0;
        if (v.wavefront)
        {
            // a pool smaller than the launch's rows takes them in bands: each band's pixels are its own
            const int band = wavefront_->rows_capacity();
            for (int row = work.first_row; row < work.first_row + work.row_count; row += band)
            {
                dev::launch_work part = work;
                part.first_row = row;
                part.row_count = std::min(band, work.first_row + work.row_count - row);
                last_bounces_ += wavefront_->render(seed, part, sums_.data(), rays_.data());
            }
        }
        else if (v.plain)
        {
            check(launch_plain_megakernel<Real>(mode_, cam_, dscene_.view(), seed, work, sums_.data(), rays_.data(),
                                                stream()),
                  "launching the kernel");
        }
        else
        {
            ensure_staging();
            ensure_far_stack();
            check(launch_megakernel<Real>(mode_, cam_, dscene_.view(), seed, work, sums_.data(), staging_.data(),
                                          staging_samples_, queue_.data(), rays_.data(), v.block, stream(),
                                          nullptr, v.waves, far_stack_.data(), far_stack_.bytes()),
                  "launching the kernel");
        }
        stop_.record(stream());
        stop_.synchronize();
        const double seconds = hip_event::seconds_between(start_, stop_);
        kernel_seconds_ += seconds;
        return seconds;
    }
};
}

void sample_numbers(uint64_t seed, const uint64_t* pixels, const uint64_t* passes, int64_t n, int draws,
                    uint32_t* words, float* floats, double* doubles, float* directions)
{
    if (n < 0 || draws < 0 || draws > 4096 || (n > 0 && (!pixels || !passes)))
    {
        throw std::invalid_argument("0 or more items with their pixels and passes, and 0 to 4096 draws");
    }
    if (n == 0)
    {
        return;
    }
    const std::size_t m = static_cast<std::size_t>(n);
    const std::size_t k = m * static_cast<std::size_t>(draws);
    device_array<uint64_t> px(m), ps(m);
    device_array<uint32_t> w(words ? k : 0);
    device_array<float> fl(floats ? k : 0), dir(directions ? 9 * m : 0);
    device_array<double> db(doubles ? k : 0);
    check(hipMemcpy(px.data(), pixels, px.bytes(), hipMemcpyHostToDevice), "copying items");
    check(hipMemcpy(ps.data(), passes, ps.bytes(), hipMemcpyHostToDevice), "copying items");
    check(launch_sample_numbers(seed, px.data(), ps.data(), n, draws, w.data(), fl.data(), db.data(), dir.data(),
                                nullptr),
          "launching the random-number probe");
    check(hipDeviceSynchronize(), "drawing the numbers");
    if (words && k)
    {
        check(hipMemcpy(words, w.data(), w.bytes(), hipMemcpyDeviceToHost), "copying words");
    }
    if (floats && k)
    {
        check(hipMemcpy(floats, fl.data(), fl.bytes(), hipMemcpyDeviceToHost), "copying floats");
    }
    if (doubles && k)
    {
        check(hipMemcpy(doubles, db.data(), db.bytes(), hipMemcpyDeviceToHost), "copying doubles");
    }
    if (directions)
    {
        check(hipMemcpy(directions, dir.data(), dir.bytes(), hipMemcpyDeviceToHost), "copying directions");
    }
}

void tone_probe(int32_t tone_map, double exposure, const float* colours, const float* dither, int64_t n,
                uint8_t* out)
{
    if (tone_map != RT_TONE_NONE && tone_map != RT_TONE_AGX && tone_map != RT_TONE_NEUTRAL)
    {
        throw std::invalid_argument("the tone probe takes RT_TONE_NONE, RT_TONE_AGX or RT_TONE_NEUTRAL");
    }
    if (!std::isfinite(exposure) || std::fabs(exposure) > 20 || n < 0)
    {
        throw std::invalid_argument("an exposure within +-20 stops and 0 or more colours");
    }
    if (n == 0)
    {
        return;
    }
    const std::size_t m = static_cast<std::size_t>(n);
    device_array<float> c(3 * m), d(dither ? m : 0);
    device_array<uint8_t> o(3 * m);
    check(hipMemcpy(c.data(), colours, c.bytes(), hipMemcpyHostToDevice), "copying colours");
    if (dither)
    {
        check(hipMemcpy(d.data(), dither, d.bytes(), hipMemcpyHostToDevice), "copying dither");
    }
    check(launch_tone_probe(c.data(), dither ? d.data() : nullptr, n, tone_map,
                            static_cast<float>(std::exp2(exposure)), o.data(), nullptr),
          "launching the tone probe");
    check(hipDeviceSynchronize(), "tone mapping");
    check(hipMemcpy(out, o.data(), o.bytes(), hipMemcpyDeviceToHost), "copying bytes");
}

void scatter_probe(int32_t lobe, float alpha, float roughness, const float color[3], int64_t n, const float* v,
                   const float* l, uint64_t seed, float* value, float* density, float* drawn, float* weight,
                   float* drawn_density)
{
    if (lobe < RT_LOBE_LAMBERT || lobe > RT_LOBE_DIELECTRIC_LEAVING)
    {
        throw std::invalid_argument("unknown scattering lobe");
    }
    if (n < 0 || !color || !(alpha >= 0.0f) || !(roughness >= 0.0f) || roughness > 1.0f ||
        (lobe >= RT_LOBE_DIELECTRIC_ENTERING && !(alpha > 0.0f)))
    {
        throw std::invalid_argument("0 or more cases, a color, alpha >= 0 and a roughness in [0, 1]");
    }
    if (n == 0)
    {
        return;
    }
    const std::size_t m = static_cast<std::size_t>(n);
    device_array<float> dv(3 * m), dl(3 * m), val(3 * m), den(m), dr(3 * m), wt(3 * m), dd(m);
    check(hipMemcpy(dv.data(), v, dv.bytes(), hipMemcpyHostToDevice), "copying directions");
    check(hipMemcpy(dl.data(), l, dl.bytes(), hipMemcpyHostToDevice), "copying directions");
    check(launch_scatter_probe(lobe, alpha, roughness, color, n, dv.data(), dl.data(), seed, val.data(), den.data(),
                               dr.data(), wt.data(), dd.data(), nullptr),
          "launching the scattering probe");
    check(hipDeviceSynchronize(), "scattering");
    check(hipMemcpy(value, val.data(), val.bytes(), hipMemcpyDeviceToHost), "copying values");
    check(hipMemcpy(density, den.data(), den.bytes(), hipMemcpyDeviceToHost), "copying densities");
    check(hipMemcpy(drawn, dr.data(), dr.bytes(), hipMemcpyDeviceToHost), "copying draws");
    check(hipMemcpy(weight, wt.data(), wt.bytes(), hipMemcpyDeviceToHost), "copying weights");
    check(hipMemcpy(drawn_density, dd.data(), dd.bytes(), hipMemcpyDeviceToHost), "copying densities");
}

namespace
{
void check_options(const rt_session_options& options)
{
    if (options.precision != RT_FLOAT32 && options.precision != RT_FLOAT64 && options.precision != RT_FLOAT64_ALL)
    {
        throw std::invalid_argument("unknown precision");
    }
    if (options.integrator < RT_INTEGRATOR_AUTO || options.integrator > RT_INTEGRATOR_PLAIN)
    {
        throw std::invalid_argument("unknown integrator");
    }
    if (options.launch_seconds < 0 || options.launch_seconds > 0.5)
    {
        throw std::invalid_argument("launch_seconds must be in (0, 0.5]");
    }
    if (options.block_size != 0 && (options.block_size < 64 || options.block_size > 256 || options.block_size % 64 != 0))
    {
        throw std::invalid_argument("the block size must be 64, 128, 192 or 256 threads");
    }
    if (options.tree < RT_TREE_DEFAULT || options.tree > RT_TREE_TWO_LEVEL)
    {
        throw std::invalid_argument("unknown tree");
    }
    if (options.nodes < RT_NODES_AUTO || options.nodes > RT_NODES_PACKED)
    {
        throw std::invalid_argument("unknown node layout");
    }
    if (options.kernel < RT_KERNEL_AUTO || options.kernel > RT_KERNEL_QUALITY)
    {
        throw std::invalid_argument("unknown kernel");
    }
    if (options.wait != RT_WAIT_SPIN && options.wait != RT_WAIT_SLEEP)
    {
        throw std::invalid_argument("unknown wait mode");
    }
    if (options.wait == RT_WAIT_SLEEP)
    {
        // an event that sleeps needs the device's waits to sleep too, else its copies still spin (as the reference
        // requires, src/gpu/renderer.h:536-540)
        unsigned flags = 0;
        check(hipGetDeviceFlags(&flags), "reading how the device waits");
        if ((flags & hipDeviceScheduleMask) != hipDeviceScheduleBlockingSync)
        {
            throw std::logic_error("a session that sleeps needs the device set to sleep first "
                                   "(rt_query_device_wait with RT_WAIT_SLEEP)");
        }
    }
}
}

std::unique_ptr<session> make_session(const rt_scene& scene, const rt_session_options& options)
{
    validate_scene(scene);
    check_options(options);
    // (the plain megakernel renders config 3 faster on float nodes, 565 against 506 Mpaths/s: its fewer waves per SIMD
    // hide less of the packed planes' unpacking; the queue megakernel, 1163 against 1038, and the wavefront take them)
    const int32_t nodes =
        options.nodes == RT_NODES_AUTO && options.integrator == RT_INTEGRATOR_PLAIN ? RT_NODES_FLOAT : options.nodes;
    return make_session(upload_scene(scene, options.precision, options.tree, nodes, options.kernel), scene.camera,
                        options);
}

std::unique_ptr<session> make_session(std::shared_ptr<const shared_scene> scene, const rt_camera_desc& camera,
                                      const rt_session_options& options)
{
    check_options(options);
    if ((options.precision != RT_FLOAT32) != (scene->f64 != nullptr))
    {
        throw std::invalid_argument("the session's precision must be the resident scene's");
    }
    if (options.precision != RT_FLOAT32)
    {
        return std::make_unique<session_impl<double>>(std::move(scene), camera, options);
    }
    return std::make_unique<session_impl<float>>(std::move(scene), camera, options);
}
}
