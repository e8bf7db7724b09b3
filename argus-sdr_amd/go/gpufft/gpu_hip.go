//go:build hip

// Package gpufft, hip build tag: forward unnormalised complex FFT on the
// device (successor of the cufft Engine, internal/fft/gpufft/gpu.go:18-91).
// Sizes: 4096 and 32768 points (the register-resident radix-16/32 kernels of
// libargus_hip.so); New returns an error for other sizes.
package gpufft

/*
#cgo LDFLAGS: -largus_hip
#include "argus_gpu.h"
*/
import "C"

import (
	"errors"
	"fmt"
	"unsafe"
)

type Engine struct {
	s     *C.gpud_surv
	n     int
	dIn   unsafe.Pointer
	dOut  unsafe.Pointer
	bytes C.size_t
}

func Available() bool {
	var n C.int
	return C.gpud_device_count(&n) == 0 && n > 0
}

func New(n int) (*Engine, error) {
	if n <= 0 {
		return nil, errors.New("invalid fft size")
	}
	if !Available() {
		return nil, errors.New("hip device not available")
	}
	e := &Engine{n: n, bytes: C.size_t(n) * 8}
	// the sample rate only scales detection-list frequencies, unused here
	if C.gpud_surv_create(C.int(n), 1, &e.s) != 0 {
		return nil, fmt.Errorf("fft size %d not supported", n)
	}
	if C.gpud_malloc(&e.dIn, e.bytes) != 0 || C.gpud_malloc(&e.dOut, e.bytes) != 0 {
		e.Close()
		return nil, errors.New("gpud_malloc failed")
	}
	return e, nil
}

func (e *Engine) Close() {
	if e == nil {
		return
	}
	if e.dIn != nil {
		_ = C.gpud_free(e.dIn)
		e.dIn = nil
	}
	if e.dOut != nil {
		_ = C.gpud_free(e.dOut)
		e.dOut = nil
	}
	if e.s != nil {
		_ = C.gpud_surv_destroy(e.s)
		e.s = nil
	}
}

// Exec: unnormalised forward transform, natural bin order (gpu.go:67-91).
func (e *Engine) Exec(in []complex64) ([]complex64, error) {
	if e == nil {
		return nil, errors.New("gpu fft not initialized")
	}
	if len(in) != e.n {
		return nil, fmt.Errorf("expected %d samples, got %d", e.n, len(in))
	}
	if C.gpud_memcpy_h2d(e.dIn, unsafe.Pointer(&in[0]), e.bytes) != 0 {
		return nil, errors.New("gpud_memcpy_h2d failed")
	}
	if C.gpud_surv_fft_device(e.s, (*C.gpud_float2)(e.dIn), C.int64_t(e.n), 1, 0, (*C.gpud_float2)(e.dOut)) != 0 {
		return nil, errors.New("gpud_surv_fft_device failed")
	}
	if C.gpud_stream_sync(C.gpud_surv_stream(e.s)) != 0 {
		return nil, errors.New("gpud_stream_sync failed")
	}
	out := make([]complex64, e.n)
	if C.gpud_memcpy_d2h(unsafe.Pointer(&out[0]), e.dOut, e.bytes) != 0 {
		return nil, errors.New("gpud_memcpy_d2h failed")
	}
	return out, nil
}
