//go:build hip

// Package rds, hip build tag: many Decoders whose state lives on the device
// (gpud_rds_batch_* of libargus_hip.so), for the per-station decode of
// cmd/sdrd/pipeline_runtime.go:1132-1142 done for every station due in one
// call.  Decode semantics are those of Decoder.Decode (rds.go:78-264).
package rds

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

// BatchDecoder holds n station decoders on the device.
type BatchDecoder struct {
	b *C.gpud_rds_batch
	n int
}

func NewBatchDecoder(stations int) (*BatchDecoder, error) {
	var b *C.gpud_rds_batch
	if rc := C.gpud_rds_batch_create(&b, C.int(stations)); rc != 0 {
		return nil, fmt.Errorf("gpud_rds_batch_create: %d", int(rc))
	}
	return &BatchDecoder{b: b, n: stations}, nil
}

func (d *BatchDecoder) Close() {
	if d != nil && d.b != nil {
		C.gpud_rds_batch_destroy(d.b)
		d.b = nil
	}
}

// Reset returns one station to a fresh Decoder.
func (d *BatchDecoder) Reset(station int) error {
	if rc := C.gpud_rds_batch_reset(d.b, C.int(station)); rc != 0 {
		return fmt.Errorf("gpud_rds_batch_reset: %d", int(rc))
	}
	return nil
}

// DecodeDevice makes one Decode call for each listed station: station
// stations[i] decodes samples [offsets[i], offsets[i+1]) of the device
// buffer samples (complex64) at rates[i].  Returns the groups found per call.
func (d *BatchDecoder) DecodeDevice(stations []int32, samples unsafe.Pointer, offsets []int64,
	rates []int32, stream unsafe.Pointer) ([]int32, error) {
	if len(offsets) != len(stations)+1 || len(rates) != len(stations) {
		return nil, errors.New("rds: offsets needs len(stations)+1 entries, rates len(stations)")
	}
	found := make([]int32, len(stations))
	if len(stations) == 0 {
		return found, nil
	}
	rc := C.gpud_rds_batch_decode_device(d.b, C.int(len(stations)), (*C.int32_t)(unsafe.Pointer(&stations[0])),
		(*C.gpud_float2)(samples), (*C.int64_t)(unsafe.Pointer(&offsets[0])),
		(*C.int32_t)(unsafe.Pointer(&rates[0])), (*C.int32_t)(unsafe.Pointer(&found[0])), C.gpud_stream_handle(stream))
	if rc != 0 {
		return nil, fmt.Errorf("gpud_rds_batch_decode_device: %d", int(rc))
	}
	return found, nil
}

// Result is the station's Result as Decoder.Decode returns it.
func (d *BatchDecoder) Result(station int) (Result, error) {
	var pi C.uint16_t
	var ps [9]C.char
	var rt [65]C.char
	if rc := C.gpud_rds_batch_state(d.b, C.int(station), &pi, &ps[0], &rt[0]); rc != 0 {
		return Result{}, fmt.Errorf("gpud_rds_batch_state: %d", int(rc))
	}
	return Result{PI: uint16(pi), PS: C.GoString(&ps[0]), RT: C.GoString(&rt[0])}, nil
}
