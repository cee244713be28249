// KSpaceFirstOrderHIP.swift -- the `.hip` device path of kspaceFirstOrder for KWaveSwift.
//
// Drop into Sources/KWave/Solver/ next to KSpaceFirstOrder.swift, add `case hip` to DeviceKind
// (Sources/KWave/KWave.swift:12-22) and, in kspaceFirstOrder (KSpaceFirstOrder.swift:94), before
// `switch grid.dim`:
//     if options.device == .hip {
//         return kspaceFirstOrderHIP(grid: grid, medium: medium, source: source, sensor: sensor, options: options)
//     }
// Package.swift gains `.systemLibrary(name: "CKWaveHIP", path: "Sources/CKWaveHIP", pkgConfig: "kwavehip")`
// (the CHDF5 pattern, Package.swift:18-22) and KWave depends on it.
//
// Not compiled in this repository (no Swift toolchain here); the C ABI it calls is the tested
// contract (tests/test_abi.py, tests/test_gpu_parity.py drive the same entry points via ctypes).

import Foundation
import MLX
import CKWaveHIP

/// Keeps host copies alive while the descriptor points into them.
private final class HostArena {
    private var floats: [UnsafeMutableBufferPointer<Float>] = []
    private var ints: [UnsafeMutableBufferPointer<Int64>] = []
    deinit {
        floats.forEach { $0.deallocate() }
        ints.forEach { $0.deallocate() }
    }
    func copy(_ a: [Float]) -> UnsafeMutablePointer<Float> {
        let b = UnsafeMutableBufferPointer<Float>.allocate(capacity: max(a.count, 1))
        _ = b.initialize(from: a)
        floats.append(b)
        return b.baseAddress!
    }
    func copy(_ a: [Int64]) -> UnsafeMutablePointer<Int64> {
        let b = UnsafeMutableBufferPointer<Int64>.allocate(capacity: max(a.count, 1))
        _ = b.initialize(from: a)
        ints.append(b)
        return b.baseAddress!
    }
}

private func hostFloats(_ a: MLXArray) -> [Float] {
    a.asType(.float32).reshaped([a.size]).asArray(Float.self)
}

private func hostIndices(_ mask: MLXArray) -> [Int64] {
    flatNonzeroIndices(mask).asArray(Int32.self).map { Int64($0) }
}

private func check(_ status: Int32, _ what: String) {
    if status != 0 {
        let msg = String(cString: kwh_last_error())
        preconditionFailure("\(what): \(msg)")
    }
}

private func field(_ a: MLXArray?, _ arena: HostArena) -> kwh_field {
    guard let a else { return kwh_field(data: nil, count: 0) }
    let h = hostFloats(a)
    return kwh_field(data: arena.copy(h), count: Int64(h.count))
}

private func sourceField(mask: MLXArray?, signal: MLXArray?, mode: SourceMode, _ arena: HostArena) -> kwh_source_field {
    guard let mask, let signal else {
        return kwh_source_field(index: nil, n_points: 0, signal: nil, rows: 0, len: 0, mode: 0, delay: nil, scale: nil)
    }
    let idx = hostIndices(mask)
    let rows = signal.ndim <= 1 ? 1 : signal.shape[0]
    let sig = hostFloats(signal)
    let m: Int32
    switch mode {
    case .dirichlet: m = 0
    case .additive: m = 1
    case .additiveNoCorrection: m = 2
    }
    return kwh_source_field(index: arena.copy(idx), n_points: Int64(idx.count), signal: arena.copy(sig),
                            rows: Int32(rows), len: Int32(sig.count / max(rows, 1)), mode: m, delay: nil, scale: nil)
}

/// The device context of the `.hip` path: opened on first use, kept for the process. A class with
/// a lock instead of a global `var`, which Swift 6 language mode rejects as unprotected shared
/// mutable state; the lock makes concurrent first uses from several threads safe.
private final class HIPContext: @unchecked Sendable {
    static let shared = HIPContext()
    private let lock = NSLock()
    private var ctx: OpaquePointer? = nil

    func get() -> OpaquePointer? {
        lock.lock()
        defer { lock.unlock() }
        if ctx == nil { check(kwh_ctx_create(0, &ctx), "kwh_ctx_create") }
        return ctx
    }
}

/// What the C step callback needs, passed through its `user` pointer (Unmanaged, retained by the
/// caller for the duration of kwh_sim_run).
private final class StepCallbacks {
    let progress: ((Int, Int) -> Void)?
    let monitor: ((Int, MLXArray) -> Void)?
    let shape: [Int]
    let count: Int
    let float64: Bool

    init(options: SimulationOptions, grid: KWaveGrid) {
        progress = options.progress
        monitor = options.fieldMonitor
        shape = grid.size
        count = grid.totalGridPoints
        float64 = options.dtype == .float64
    }

    /// One call per step t (in order); pHost is the pressure after step t in the sim's precision
    /// when the field is due, else nil.
    func step(_ t: Int, _ nt: Int, _ pHost: UnsafeRawPointer?) {
        if let monitor, let pHost {
            let p: MLXArray
            if float64 {
                let b = UnsafeBufferPointer(start: pHost.assumingMemoryBound(to: Double.self), count: count)
                p = MLXArray(converting: Array(b)).reshaped(shape)
            } else {
                let b = UnsafeBufferPointer(start: pHost.assumingMemoryBound(to: Float.self), count: count)
                p = MLXArray(Array(b)).reshaped(shape)
            }
            monitor(t, p)
        }
        progress?(t, nt)
    }
}

/// Flattens the API values into a descriptor whose pointers live in `arena`.
private func makeDescriptor(grid: KWaveGrid, medium: KWaveMedium, source: KWaveSource, sensor: KWaveSensor,
                            options: SimulationOptions, arena: HostArena) -> (kwh_sim_desc, Int) {
    precondition(grid.nt > 0, "call grid.makeTime(...) before running the solver")
    precondition(source.p0 != nil || source.pMask != nil || source.uMask != nil,
                 "\(grid.dim)D solver requires source.p0 or a time-varying p/u source")
    var d = kwh_sim_desc()
    d.abi_version = Int32(KWH_ABI_VERSION)
    d.dim = Int32(grid.dim)
    d.n = (Int32(grid.nx), Int32(grid.ny), Int32(grid.nz))
    d.d = (grid.dx, grid.dy, grid.dz)
    d.dt = grid.dt
    d.nt = Int32(grid.nt)
    d.dtype = options.dtype == .float64 ? 1 : 0
    switch options.pmlSize { case let .uniform(s): d.pml_size = Int32(s) }
    d.pml_alpha = options.pmlAlpha
    d.smooth_p0 = options.smoothP0 ? 1 : 0
    d.smooth_c0 = options.smoothC0 ? 1 : 0
    d.smooth_rho0 = options.smoothRho0 ? 1 : 0
    d.sound_speed = field(medium.soundSpeed, arena)
    d.density = field(medium.density, arena)
    d.alpha_coeff = field(medium.alphaCoeff, arena)
    d.alpha_power = medium.alphaPower ?? 1.5
    switch medium.alphaMode {
    case .noAbsorption: d.alpha_mode = 0
    case .powerLaw: d.alpha_mode = 1
    case .noDispersion: d.alpha_mode = 2
    case .stokes: d.alpha_mode = 3
    }
    d.bona = field(medium.bOnA, arena)
    d.p0 = field(source.p0, arena)
    d.p_source = sourceField(mask: source.pMask, signal: source.p, mode: source.pMode, arena)
    d.u_source = (sourceField(mask: source.uMask, signal: source.ux, mode: source.uMode, arena),
                  sourceField(mask: source.uMask, signal: source.uy, mode: source.uMode, arena),
                  sourceField(mask: source.uMask, signal: source.uz, mode: source.uMode, arena))
    var nSensor = 0
    if let mask = sensor.mask {
        if mask.ndim == 2 && mask.dim(0) == grid.dim {              // Recording.swift:50-53
            let pts = hostFloats(mask)
            d.sensor_cartesian = UnsafePointer(arena.copy(pts))
            d.n_cartesian = Int64(mask.dim(1))
            nSensor = mask.dim(1)
        } else {
            let idx = hostIndices(mask)
            d.sensor_index = UnsafePointer(arena.copy(idx))
            d.n_sensor = Int64(idx.count)
            nSensor = idx.count
        }
        if let ang = sensor.directivityAngle {
            d.directivity_angle = UnsafePointer(arena.copy(hostFloats(ang)))
            d.directivity_size = sensor.directivitySize ?? 0
            d.directivity_pattern = sensor.directivityPattern == .gradient ? 1 : 0
        }
    }
    d.record = UInt32(sensor.record.rawValue)
    return (d, nSensor)
}

private func createSim(_ d: inout kwh_sim_desc, axisymmetric: Bool = false) -> OpaquePointer? {
    let ctx = HIPContext.shared.get()
    var sim: OpaquePointer? = nil
    if axisymmetric {
        check(kwh_sim_create_axisymmetric(ctx, &d, &sim), "kwh_sim_create_axisymmetric")
    } else {
        check(kwh_sim_create(ctx, &d, &sim), "kwh_sim_create")
    }
    return sim
}

/// `axisymmetric: true` runs kspaceFirstOrderAS (KSpaceFirstOrderAS.swift:17-309) on the device:
/// x axial, y radial; the output carries p and pFinal.
func kspaceFirstOrderHIP(grid: KWaveGrid, medium: KWaveMedium, source: KWaveSource,
                         sensor: KWaveSensor, options: SimulationOptions,
                         axisymmetric: Bool = false) -> SimulationOutput {
    let arena = HostArena()
    var (d, nSensor) = makeDescriptor(grid: grid, medium: medium, source: source, sensor: sensor,
                                      options: options, arena: arena)
    let sim = createSim(&d, axisymmetric: axisymmetric)
    defer { kwh_sim_destroy(sim) }

    if options.progress == nil && options.fieldMonitor == nil {
        check(kwh_sim_run(sim, 0, Int32(grid.nt), nil, 0, nil), "kwh_sim_run")
    } else {
        // progress(t, nt) every step and fieldMonitor(t, p) every fieldMonitorInterval steps, as in
        // the reference (:337-340), delivered by the library's step callback: steps run in chunks
        // between device syncs, and the field is copied to the host only on the steps it is due.
        let callbacks = StepCallbacks(options: options, grid: grid)
        let user = Unmanaged.passRetained(callbacks).toOpaque()
        defer { Unmanaged<StepCallbacks>.fromOpaque(user).release() }
        let interval: Int32 = options.fieldMonitor == nil ? 0 : Int32(max(options.fieldMonitorInterval, 1))
        check(kwh_sim_run(sim, 0, Int32(grid.nt), { t, nt, pHost, user in
            guard let user else { return }
            Unmanaged<StepCallbacks>.fromOpaque(user).takeUnretainedValue().step(Int(t), Int(nt), pHost)
        }, interval, user), "kwh_sim_run")
    }

    func read(_ which: kwh_output) -> MLXArray? {
        var shape: (Int64, Int64, Int64) = (1, 1, 1)
        var nd: Int32 = 0
        let ok = withUnsafeMutablePointer(to: &shape) { p in
            p.withMemoryRebound(to: Int64.self, capacity: 3) { kwh_sim_output_shape(sim, Int32(which.rawValue), $0, &nd) }
        }
        guard ok == 0 else { return nil }
        let dims = [Int(shape.0), Int(shape.1), Int(shape.2)].prefix(Int(nd)).map { $0 }
        let count = dims.reduce(1, *)
        if options.dtype == .float64 {
            var h = [Double](repeating: 0, count: count)
            h.withUnsafeMutableBytes { b in check(kwh_sim_read_output(sim, Int32(which.rawValue), b.baseAddress, b.count), "read") }
            return MLXArray(converting: h).reshaped(dims)
        }
        var h = [Float](repeating: 0, count: count)
        h.withUnsafeMutableBytes { b in check(kwh_sim_read_output(sim, Int32(which.rawValue), b.baseAddress, b.count), "read") }
        return MLXArray(h).reshaped(dims)
    }

    _ = nSensor
    var out = SimulationOutput()
    out.pFinal = read(KWH_OUT_P_FINAL)
    out.p = read(KWH_OUT_P); out.pMax = read(KWH_OUT_P_MAX); out.pMin = read(KWH_OUT_P_MIN); out.pRms = read(KWH_OUT_P_RMS)
    out.ux = read(KWH_OUT_UX); out.uy = read(KWH_OUT_UY); out.uz = read(KWH_OUT_UZ)
    out.uxMax = read(KWH_OUT_UX_MAX); out.uyMax = read(KWH_OUT_UY_MAX); out.uzMax = read(KWH_OUT_UZ_MAX)
    out.uxRms = read(KWH_OUT_UX_RMS); out.uyRms = read(KWH_OUT_UY_RMS); out.uzRms = read(KWH_OUT_UZ_RMS)
    out.uxFinal = read(KWH_OUT_UX_FINAL); out.uyFinal = read(KWH_OUT_UY_FINAL); out.uzFinal = read(KWH_OUT_UZ_FINAL)
    out.ixAvg = read(KWH_OUT_IX_AVG); out.iyAvg = read(KWH_OUT_IY_AVG); out.izAvg = read(KWH_OUT_IZ_AVG)
    return out
}

/// Forward / VJP pair with the contract of a torch.autograd.Function (no counterpart in the
/// reference): `forward()` runs the loop saving checkpoints and returns the recorded pressure
/// [nSensor, nt]; `vjp(dL/dp)` returns (dL/dp0, dL/dc) for the lossless linear p0 problem, at
/// fixed c_ref = max(c). Float32 path (options.dtype .float32).
public final class KSpaceFirstOrderHIPVJP {
    private let arena = HostArena()
    private let sim: OpaquePointer?
    private let grid: KWaveGrid
    private let nSensor: Int
    private let checkpointEvery: Int32

    public init(grid: KWaveGrid, medium: KWaveMedium, source: KWaveSource, sensor: KWaveSensor,
                options: SimulationOptions, checkpointEvery: Int = 32) {
        var (d, n) = makeDescriptor(grid: grid, medium: medium, source: source, sensor: sensor,
                                    options: options, arena: arena)
        precondition(d.dtype == 0, "the VJP pair runs in float32")
        self.sim = createSim(&d)
        self.grid = grid
        self.nSensor = n
        self.checkpointEvery = Int32(checkpointEvery)
    }
    deinit { kwh_sim_destroy(sim) }

    public func forward() -> MLXArray {
        check(kwh_sim_set_checkpointing(sim, checkpointEvery), "kwh_sim_set_checkpointing")
        check(kwh_sim_run(sim, 0, Int32(grid.nt), nil, 0, nil), "kwh_sim_run")
        var h = [Float](repeating: 0, count: nSensor * grid.nt)
        h.withUnsafeMutableBytes { b in
            check(kwh_sim_read_output(sim, Int32(KWH_OUT_P.rawValue), b.baseAddress, b.count), "read p")
        }
        return MLXArray(h).reshaped([nSensor, grid.nt])
    }

    public func vjp(_ dLdp: MLXArray) -> (p0: MLXArray, c: MLXArray) {
        let g = dLdp.asType(.float32).reshaped([nSensor * grid.nt]).asArray(Float.self)
        var gp0 = [Float](repeating: 0, count: grid.totalGridPoints)
        var gc = [Float](repeating: 0, count: grid.totalGridPoints)
        g.withUnsafeBytes { gb in
            gp0.withUnsafeMutableBytes { pb in
                gc.withUnsafeMutableBytes { cb in
                    check(kwh_sim_vjp(sim, gb.baseAddress, checkpointEvery, pb.baseAddress, cb.baseAddress),
                          "kwh_sim_vjp")
                }
            }
        }
        return (MLXArray(gp0).reshaped(grid.size), MLXArray(gc).reshaped(grid.size))
    }
}
