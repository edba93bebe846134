// Exception-to-status mapping shared by the extern "C" entry points: the
// reference's exceptions (invalid_argument, length_error, logic_error,
// runtime_error) become distinct status codes, with the message kept per
// thread for c8mi_last_error().
#pragma once

#include <exception>
#include <new>
#include <stdexcept>

#include "runtime.hpp"

namespace c8mi {

template <class F>
int guarded(F&& body) {
  try {
    body();
    set_last_error("");
    return kOk;
  } catch (OverflowError const& e) {
    set_last_error(e.what());
    return kOverflowError;
  } catch (DeviceError const& e) {
    set_last_error(e.what());
    return kDeviceError;
  } catch (std::invalid_argument const& e) {
    set_last_error(e.what());
    return kInvalidArgument;
  } catch (std::length_error const& e) {
    set_last_error(e.what());
    return kLengthError;
  } catch (std::logic_error const& e) {
    set_last_error(e.what());
    return kLogicError;
  } catch (std::bad_alloc const&) {
    set_last_error("host allocation failed");
    return kRuntimeError;
  } catch (std::exception const& e) {
    set_last_error(e.what());
    return kRuntimeError;
  } catch (...) {
    set_last_error("unknown error");
    return kRuntimeError;
  }
}

}  // namespace c8mi
