// Host-side session object: one reconstruction (volume, live-frame maps,
// model maps, pose state) with its own HIP stream.  The public opaque types
// dm_volume and dm_session both point at this object.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "../../include/dynamap.h"
#include "../../include/dynamap_session.h"
#include "dm_internal.h"

namespace dm {

void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define DM_HIP(call)                                 \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return hip_fail(e_, #call); \
  } while (0)

// propagate a DM_* status
#define DM_TRY(x)                  \
  do {                             \
    int rc_ = (x);                 \
    if (rc_ != DM_OK) return rc_;  \
  } while (0)

struct Sess {
  dm_tsdf_props props{};
  dm_icp_props icp{};
  uint32_t max_rows = 0, max_cols = 0;
  DmView h{};                    // host copy of the view
  DmView *d_view = nullptr;      // device copy (one-entry table)
  int *d_active = nullptr;       // device {0}
  hipStream_t stream = nullptr;
  int device = 0;                // HIP device the state lives on (current at init)
  std::vector<void *> allocs;    // device allocations owned by the session
  float *d_depth = nullptr;      // frame staging
  uint8_t *d_rgb = nullptr;
  uint8_t *h_pinned = nullptr;   // pinned staging for host inputs
  size_t pinned_bytes = 0;
  float *d_ray_depth = nullptr;  // raycast outputs for host readback
  uint8_t *d_ray_rgb = nullptr;
  float *d_depth_f = nullptr;    // bilateral-filtered depth for tracking (when enabled)
  int filt_r = 0;                // filter radius, 0 = off
  int icp_mfma = 0;              // ICP linearisation on the matrix cores (dm_session_set_icp_reduce)
  float filt_sd = 0.f, filt_ss = 0.f;
  size_t hbm_bytes = 0;          // device bytes owned
  int frames = 0;

  int init(const dm_tsdf_props &p, uint32_t rows, uint32_t cols);
  void release();
  template <class T> int dalloc(T **ptr, size_t count, bool zero = true);
  int push_view();               // copy h to d_view (async on stream)
  Launch launch() const;
  int set_sensor(const dm_sensor *s);
  int set_pose_f(const float pose[16]);   // T and Tinv (float), T in double too
  int stage_frame(const float *host_depth, const uint8_t *host_rgb);
};

// bytes of device memory a session with these parameters needs
size_t session_bytes(const dm_tsdf_props &p, uint32_t rows, uint32_t cols);
int validate_props(const dm_tsdf_props &p);

}  // namespace dm
