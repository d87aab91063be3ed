// capi_util.h — status codes and the per-thread error message of the C ABI.
#pragma once

#include <exception>
#include <stdexcept>
#include <string>

#include "../../include/rustracer_hip.h"

struct rt_scene;

namespace rt {

struct Scene;

class ApiError : public std::runtime_error {
  public:
    ApiError(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

void set_last_error(const std::string &msg);

// Run `f`, translating exceptions into a status code and rt_last_error().
template <class F>
int api_call(F &&f) {
    try {
        f();
        set_last_error("");
        return RT_OK;
    } catch (const ApiError &e) {
        set_last_error(e.what());
        return e.code();
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return RT_ERR_CAPACITY;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return RT_ERR_INVALID;
    }
}

const Scene &scene_of(const rt_scene *s);
Scene &scene_of(rt_scene *s);

}  // namespace rt
