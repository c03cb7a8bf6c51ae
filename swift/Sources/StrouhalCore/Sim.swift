// StrouhalCore's solver API (Sources/StrouhalCore/Sim.swift:5-459 of the
// reference) over the MI355X library. Names, argument labels and error
// behaviour follow the Metal-backed classes so that Cases.swift,
// CaseSetups.swift and Ladder.swift compile against it; the differences are
// that probeMoments() returns a copy, and that encode(steps:on:) (the Metal
// render loop's entry) does not exist.
import CStrouhalHIP
import Foundation

public enum StrouhalError: Error, CustomStringConvertible {
    case noDevice
    case message(String)
    public var description: String {
        switch self {
        case .noDevice: return "no HIP device"
        case .message(let m): return m
        }
    }
}

@inline(__always) func check(_ rc: Int32) throws {
    guard rc == Int32(STRH_OK) else {
        let msg = String(cString: strh_last_error())
        throw rc == Int32(STRH_ERR_NO_DEVICE) ? StrouhalError.noDevice : StrouhalError.message(msg)
    }
}

public enum Precision: Int32 {                    // Sim.swift:5-8 (+ bf16)
    case fp32 = 0, fp16s = 1, bf16 = 2
    public var ddfBytes: Int { self == .fp32 ? 4 : 2 }
}

public enum Cell: UInt8 {                         // Sim.swift:46-52
    case fluid = 0, solid = 1, lid = 2, inflow = 3, outflow = 4
}

public final class GPU {                          // Sim.swift:60-78
    let ctx: OpaquePointer
    public init(device: Int32 = 0) throws {
        var c: OpaquePointer?
        try check(strh_ctx_create(device, &c))
        ctx = c!
    }
    deinit { strh_ctx_destroy(ctx) }
    public var name: String {
        var buf = [CChar](repeating: 0, count: 256)
        strh_device_name(ctx, &buf, buf.count)
        return String(cString: buf)
    }
}

public final class Simulation {                   // Sim.swift:153-459
    public let gpu: GPU
    public let precision: Precision
    let lat: OpaquePointer
    public let nx: Int, ny: Int, nz: Int
    public var cells: Int { nx * ny * nz }

    public static let cx: [Int] = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]
    public static let cy: [Int] = [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1]
    public static let cz: [Int] = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1]

    /// tau -> (omega+, omega-) via the TRT magic parameter Lambda (Sim.swift:188-193).
    public static func trtOmegas(tau: Double, lambda: Double?) -> (Float, Float) {
        let wp = 1.0 / tau
        guard let lambda else { return (Float(wp), Float(wp)) }
        return (Float(wp), Float(1.0 / (0.5 + lambda / (tau - 0.5))))
    }

    public init(gpu: GPU, precision: Precision = .fp32, nx: Int, ny: Int, nz: Int,
                omega: Float, omegaMinus: Float? = nil,
                lid: SIMD3<Float> = .zero, uin: Float = 0, rampSteps: Int = 0,
                force: SIMD3<Float> = .zero, cSmago: Float = 0,
                wantsForces: Bool = false,
                flags: (Int, Int, Int) -> Cell) throws {
        self.gpu = gpu
        self.precision = precision
        self.nx = nx; self.ny = ny; self.nz = nz
        var fl = [UInt8](repeating: 0, count: nx * ny * nz)
        for z in 0..<nz { for y in 0..<ny { for x in 0..<nx {
            fl[(z * ny + y) * nx + x] = flags(x, y, z).rawValue
        }}}
        var d = strh_lattice_desc(precision: precision.rawValue,
                                  nx: Int32(nx), ny: Int32(ny), nz: Int32(nz),
                                  omega: omega, omega_minus: omegaMinus ?? -1,
                                  lid: (lid.x, lid.y, lid.z), uin: uin,
                                  ramp_steps: Int64(rampSteps),
                                  force: (force.x, force.y, force.z), c_smago: cSmago,
                                  wants_forces: wantsForces ? 1 : 0)
        var h: OpaquePointer?
        try check(strh_lattice_create(gpu.ctx, &d, fl, &h))
        lat = h!
        self.omega = omega
        self.omegaMinus = omegaMinus ?? omega
        self.lidVel = lid
        self.uinTarget = uin
        self.rampSteps = rampSteps
        self.force = force
        self.cSmago = cSmago
    }
    deinit { strh_lattice_destroy(lat) }

    // ---- mutable state, Sim.swift:164-172 ---------------------------------
    public var omega: Float { didSet { strh_set_omega(lat, omega, omegaMinus) } }
    public var omegaMinus: Float { didSet { strh_set_omega(lat, omega, omegaMinus) } }
    public var lidVel: SIMD3<Float> { didSet { strh_set_lid_velocity(lat, lidVel.x, lidVel.y, lidVel.z) } }
    public var uinTarget: Float { didSet { strh_set_uin(lat, uinTarget) } }
    public var rampSteps: Int { didSet { strh_set_ramp_steps(lat, Int64(rampSteps)) } }
    public var force: SIMD3<Float> { didSet { strh_set_force(lat, force.x, force.y, force.z) } }
    public var cSmago: Float { didSet { strh_set_c_smago(lat, cSmago) } }
    public var inflowUniform = false { didSet { strh_set_inflow_uniform(lat, inflowUniform ? 1 : 0) } }
    public var sponge: (x0: Float, width: Float, tau: Float)? = nil {
        didSet {
            strh_set_sponge(lat, sponge == nil ? 0 : 1, sponge?.x0 ?? 0, sponge?.width ?? 1, sponge?.tau ?? 1)
        }
    }

    public var stepsDone: Int { var s: Int64 = 0; strh_steps_done(lat, &s); return Int(s) }
    public var gpuSeconds: Double { var s = 0.0; strh_gpu_seconds(lat, &s); return s }

    // ---- setup --------------------------------------------------------------
    public func setSolidFractions(_ eps: [Float]) throws {          // Sim.swift:413-421
        precondition(eps.count == cells)
        try check(strh_set_solid_fractions(lat, eps))
    }
    /// Brinkman porosity per cell (new; nil removes it).
    public func setPorosity(_ alpha: [Float]?) throws {
        if let alpha {
            precondition(alpha.count == cells)
            try check(strh_set_porosity(lat, alpha))
        } else {
            try check(strh_set_porosity(lat, nil))
        }
    }
    public func initField(mode: UInt32, amplitude: Float) throws {   // Sim.swift:424-439
        try check(strh_init_field(lat, mode, amplitude))
    }

    // ---- stepping and observables ------------------------------------------
    public func run(steps count: Int) throws { try check(strh_run(lat, Int64(count))) }   // Sim.swift:288-331

    public func probeMoments() throws -> [SIMD4<Float>] {            // Sim.swift:334-353
        var out = [SIMD4<Float>](repeating: .zero, count: cells)
        try out.withUnsafeMutableBytes {
            try check(strh_probe_moments(lat, $0.baseAddress!.assumingMemoryBound(to: Float.self)))
        }
        return out
    }

    public func probeForce(xRange: ClosedRange<Int>, yRange: ClosedRange<Int>) throws -> SIMD3<Double> {
        var f = [Double](repeating: 0, count: 3)                    // Sim.swift:359-372
        try check(strh_probe_force(lat, Int32(xRange.lowerBound), Int32(xRange.upperBound),
                                   Int32(yRange.lowerBound), Int32(yRange.upperBound), &f))
        return SIMD3(f[0], f[1], f[2])
    }

    public var stateDigest: String {                                 // Sim.swift:441-444
        var d = [UInt8](repeating: 0, count: 32)
        strh_state_digest(lat, &d)
        return d.map { String(format: "%02x", $0) }.joined()
    }

    public func massSum() -> Double { var m = 0.0; strh_mass_sum(lat, &m); return m }   // Sim.swift:447-458
    public func kineticEnergy() throws -> Double { var e = 0.0; try check(strh_kinetic_energy(lat, &e)); return e }
    public func dissipation() throws -> Double { var e = 0.0; try check(strh_dissipation(lat, &e)); return e }

    // ---- raw state ----------------------------------------------------------
    public func downloadDDF() throws -> [UInt8] {
        var n = 0
        try check(strh_ddf_bytes(lat, &n))
        var out = [UInt8](repeating: 0, count: n)
        try check(strh_download_ddf(lat, &out, n))
        return out
    }
    public func uploadDDF(_ bytes: [UInt8]) throws { try check(strh_upload_ddf(lat, bytes, bytes.count)) }
    public func snapshot() throws { try check(strh_snapshot(lat)) }
    public func restore() throws { try check(strh_restore(lat)) }
    public func saveState(to path: String) throws { try check(strh_save_state(lat, path)) }
    public func loadState(from path: String) throws { try check(strh_load_state(lat, path)) }

    // ---- differentiation (new) ----------------------------------------------
    public struct Objective {
        public var inletX: Int32, outletX: Int32
        public var wP: Double, wD: Double
        public var spinup: Int, avg: Int
        public var wF: Double
        public init(inletX: Int32, outletX: Int32, wP: Double = 1, wD: Double = 1, spinup: Int, avg: Int,
                    wF: Double = 0) {
            self.inletX = inletX; self.outletX = outletX; self.wP = wP; self.wD = wD
            self.spinup = spinup; self.avg = avg; self.wF = wF
        }
        var c: strh_objective {
            strh_objective(inlet_x: inletX, outlet_x: outletX, w_p: wP, w_d: wD,
                           spinup: Int64(spinup), avg: Int64(avg), w_f: wF)
        }
    }

    public func objective(_ o: Objective) throws -> Double {
        var oc = o.c
        var j = 0.0
        try check(strh_objective_eval(lat, &oc, &j))
        return j
    }

    /// J, dJ/dalpha and dJ/deps per cell and dJ/d(uin target); checkpointEvery
    /// 0 = sqrt(T), -1 = as long as the recording fits half the free device memory.
    public func vjp(_ o: Objective, checkpointEvery: Int = -1) throws
        -> (j: Double, gradAlpha: [Double], gradEps: [Double], gradUin: Double) {
        var oc = o.c
        var j = 0.0, gu = 0.0
        var ga = [Double](repeating: 0, count: cells)
        var ge = [Double](repeating: 0, count: cells)
        var info = strh_vjp_info()
        try check(strh_vjp(lat, &oc, Int64(checkpointEvery), &j, &ga, &ge, &gu, &info))
        return (j, ga, ge, gu)
    }

    /// vjp with two checkpoint levels (strh_vjp_nested): outer segments of
    /// outerEvery steps replayed once to checkpoint inner segments of
    /// innerEvery steps, for horizons whose one-level plan does not fit.
    public func vjpNested(_ o: Objective, outerEvery: Int, innerEvery: Int) throws
        -> (j: Double, gradAlpha: [Double], gradEps: [Double], gradUin: Double) {
        var oc = o.c
        var j = 0.0, gu = 0.0
        var ga = [Double](repeating: 0, count: cells)
        var ge = [Double](repeating: 0, count: cells)
        var info = strh_vjp_info()
        try check(strh_vjp_nested(lat, &oc, Int64(outerEvery), Int64(innerEvery), &j, &ga, &ge, &gu, &info))
        return (j, ga, ge, gu)
    }
}

/// Many lattices advanced together (replaces Ladder's sequential rung loop,
/// Ladder.swift:54-121).
public final class Batch {
    let h: OpaquePointer
    public let sims: [Simulation]
    public init(_ sims: [Simulation]) throws {
        self.sims = sims
        var lats: [OpaquePointer?] = sims.map { $0.lat }
        var b: OpaquePointer?
        try check(strh_batch_create(sims[0].gpu.ctx, &lats, Int32(sims.count), &b))
        h = b!
    }
    deinit { strh_batch_destroy(h) }
    public func setProbe(_ i: Int, firstStep: Int, period: Int, count: Int,
                         xRange: ClosedRange<Int>, yRange: ClosedRange<Int>) throws {
        try check(strh_batch_set_probe(h, Int32(i), Int64(firstStep), Int64(period), Int64(count),
                                       Int32(xRange.lowerBound), Int32(xRange.upperBound),
                                       Int32(yRange.lowerBound), Int32(yRange.upperBound)))
    }
    public func run(steps: [Int]) throws {
        let s = steps.map { Int64($0) }
        try check(strh_batch_run(h, s))
    }
    public func probeSeries(_ i: Int, capacity: Int) throws -> [SIMD3<Double>] {
        var out = [Double](repeating: 0, count: 3 * max(capacity, 1))
        var taken: Int64 = 0
        try check(strh_batch_probe_series(h, Int32(i), &out, Int64(capacity), &taken))
        return (0..<Int(taken)).map { SIMD3(out[3 * $0], out[3 * $0 + 1], out[3 * $0 + 2]) }
    }
    public func autotune() throws -> Int { var bs: Int32 = 0; try check(strh_batch_autotune(h, &bs)); return Int(bs) }
    public func setGraphSteps(_ n: Int) throws { try check(strh_batch_set_graph_steps(h, Int32(n))) }
    /// 0: stream every step through memory; 1: hold lattices one cell deep in LDS where the
    /// active ones fit on chip in one launch; 2: everywhere, time-sliced; 3 (default): as 1,
    /// and where groups would take turns, the faster of the two measured per segment.
    public func setResident(_ mode: Int) throws { try check(strh_batch_set_resident(h, Int32(mode))) }
    public func setResidentTimeout(seconds: Double) throws {
        try check(strh_batch_set_resident_timeout(h, seconds))
    }
}
