/* System-library shim for the Swift package: exposes the libkwavehip C ABI to Swift. */
#ifndef CKWAVEHIP_SHIM_H
#define CKWAVEHIP_SHIM_H
#include "kwave_hip.h"
#endif
