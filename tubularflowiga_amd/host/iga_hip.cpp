// iga_hip: the MI355X drop-in for `iga_cuda navier-stokes | mesh-check | device-info`
// (solvers/cuda/src/iga_cuda.cu:447-498 options, :734-1151 driver, :1157-1194 main).
// Same command line, input files, output files and stdout key=value records;
// the solve runs in libtfi_hip through the C-ABI of include/tfi.h.  Options the
// reference does not have select the MI355X solver path (preconditioner,
// precision, warm start, tangent) and default to the reproducing settings.
#include "case_input.hpp"
#include "outlets.hpp"
#include "solve_input.hpp"
#include "transport_input.hpp"
#include "tfi.h"
#ifdef TFI_HAVE_HDF5
#include "bezier_output.hpp"
#endif

#include <chrono>
#include <functional>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

namespace {

using namespace tfi_host;
using Clock = std::chrono::steady_clock;

double Seconds(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

struct Ctx {
	tfi_ctx* ctx = nullptr;
	explicit Ctx(int device)
	{
		if (tfi_ctx_create(device, &ctx) != TFI_OK) {
			const std::string msg = tfi_last_error(nullptr);
			throw std::runtime_error("cannot create the HIP context: " + msg);
		}
	}
	~Ctx() { tfi_ctx_destroy(ctx); }
	void Check(int status, const char* what) const
	{
		if (status != TFI_OK) throw std::runtime_error(std::string(what) + ": " + tfi_last_error(ctx));
	}
};

struct Mesh {
	tfi_mesh* mesh = nullptr;
	~Mesh() { tfi_mesh_destroy(mesh); }
};

struct Flow {
	tfi_flow* flow = nullptr;
	~Flow() { tfi_flow_destroy(flow); }
};

double Gib(double bytes) { return bytes/(1024.0*1024.0*1024.0); }

void PrintDevice(Ctx& c)
{
	char buf[512];
	c.Check(tfi_device_info(c.ctx, buf, sizeof buf), "device info");
	std::cout << buf << '\n';
}

void PrintMesh(const FlatDatabase& db, const tfi_mesh_stats* st)
{
	std::cout << "mesh nodes=" << db.nodes << " elements=" << db.elements() << " basis_entries=" << db.connectivity.size()
		<< " extraction_nonzeros=" << db.extraction_values.size() << " max_basis=" << db.maximum_basis;
	if (st)
		std::cout << " block_nonzeros=" << st->block_nonzeros
			<< " mean_block_row=" << static_cast<double>(st->block_nonzeros)/static_cast<double>(db.nodes);
	std::cout << '\n';
}

void CreateMesh(Ctx& c, const FlatDatabase& db, Mesh& m)
{
	c.Check(tfi_mesh_create(c.ctx, static_cast<int>(db.nodes), db.elements(), db.element_offsets.data(),
		db.connectivity.data(), db.extraction_offsets.data(), db.extraction_columns.data(), db.extraction_values.data(),
		db.bezier.data(), &m.mesh), "mesh upload");
}

int MeshCheck(Ctx& c, const std::string& path)
{
	const FlatDatabase db = ReadDatabase(path);
	Mesh m;
	CreateMesh(c, db, m);
	tfi_mesh_stats st{};
	c.Check(tfi_mesh_stats_get(m.mesh, &st), "mesh statistics");
	PrintMesh(db, &st);
	std::cout << std::setprecision(17) << "geometry min_detJ=" << st.minimum_determinant
		<< " bad_samples=" << st.bad_samples << " colors=" << st.colors << '\n';
	return st.bad_samples ? 1 : 0;
}

struct Options {
	fs::path database, case_dir, output, checkpoint, restart;
	std::string system;
	int maximum_newton = 12, output_every = 0, checkpoint_every = 0, stop_after_step = 0;
	double nonlinear_rtol = 1e-5, nonlinear_atol = 1e-10, mass_rtol = 1e-3;
	std::string visualization = "auto";
	// MI355X solver path
	int preconditioner = TFI_PC_AMG, precision = TFI_FP64, warm_start = TFI_WARM_PREVIOUS;
	int tangent = TFI_TANGENT_EXACT, smoother = TFI_SMOOTH_AUTO, restart_length = 200;
	int linear_maximum_iterations = 5000;
	double linear_rtol = 1e-8;
	int chord_polish = 1, train_steps = 0, progress = 0, chebyshev_degree = 0, half_smoother = 1;
};

int PositiveInt(const std::string& t, const std::string& opt)
{
	std::size_t used = 0;
	int v = 0;
	try {
		v = std::stoi(t, &used);
	} catch (const std::exception&) {
		throw std::runtime_error(opt + " requires a positive integer");
	}
	if (used != t.size() || v <= 0) throw std::runtime_error(opt + " requires a positive integer");
	return v;
}

double PositiveDouble(const std::string& t, const std::string& opt)
{
	std::size_t used = 0;
	double v = 0.0;
	try {
		v = std::stod(t, &used);
	} catch (const std::exception&) {
		throw std::runtime_error(opt + " requires a finite positive value");
	}
	if (used != t.size() || !std::isfinite(v) || v <= 0.0) throw std::runtime_error(opt + " requires a finite positive value");
	return v;
}

int Choice(const std::string& v, const std::string& opt, std::initializer_list<std::pair<const char*, int>> table)
{
	std::string names;
	for (const auto& t : table) {
		if (v == t.first) return t.second;
		names += std::string(names.empty() ? "" : "|") + t.first;
	}
	throw std::runtime_error(opt + " must be " + names);
}

const char* kUsage =
	"usage: iga_hip navier-stokes DATABASE.ntiga CASE_DIR [MAX_NEWTON] [OUTPUT] "
	"[--system NAME] [--max-newton N] [--output PATH] [--output-every N] "
	"[--checkpoint PREFIX] [--checkpoint-every N] [--restart PREFIX] "
	"[--stop-after-step N] [--nonlinear-rtol R] [--nonlinear-atol A] "
	"[--mass-rtol R] [--visualization-format auto|vtu|vtkhdf] "
	"[--preconditioner amg|bjacobi] [--precision fp64|mixed] [--warm-start previous|extrapolate|extrapolate2|extrapolate3|extrapolate4|mpe|learned] "
	"[--tangent exact|legacy] [--amg-smoother auto|chebyshev|gauss-seidel|jacobi] [--linear-rtol R] [--restart-length N] "
	"[--linear-max-iterations N] [--chord-polish on|off] [--train-steps N] [--progress on|off] "
	"[--amg-chebyshev-degree N] [--amg-half-smoother on|off]";

// ParseCudaFlowOptions (iga_cuda.cu:447-498) plus the solver-path options
Options ParseOptions(int argc, char** argv)
{
	if (argc < 4) throw std::runtime_error(kUsage);
	Options o;
	o.database = argv[2];
	o.case_dir = argv[3];
	int positional = 0;
	for (int i = 4; i < argc; ++i) {
		const std::string arg(argv[i]);
		if (arg.rfind("--", 0) != 0) {
			if (positional == 0) o.maximum_newton = PositiveInt(arg, "MAX_NEWTON");
			else if (positional == 1) o.output = arg;
			else throw std::runtime_error("too many positional arguments");
			++positional;
			continue;
		}
		if (i + 1 >= argc) throw std::runtime_error(arg + " requires a value");
		const std::string v(argv[++i]);
		if (arg == "--system") o.system = v;
		else if (arg == "--max-newton") o.maximum_newton = PositiveInt(v, arg);
		else if (arg == "--output") o.output = v;
		else if (arg == "--output-every") o.output_every = PositiveInt(v, arg);
		else if (arg == "--checkpoint") o.checkpoint = v;
		else if (arg == "--checkpoint-every") o.checkpoint_every = PositiveInt(v, arg);
		else if (arg == "--restart") o.restart = v;
		else if (arg == "--stop-after-step") o.stop_after_step = PositiveInt(v, arg);
		else if (arg == "--nonlinear-rtol") o.nonlinear_rtol = PositiveDouble(v, arg);
		else if (arg == "--nonlinear-atol") o.nonlinear_atol = PositiveDouble(v, arg);
		else if (arg == "--mass-rtol") o.mass_rtol = PositiveDouble(v, arg);
		else if (arg == "--visualization-format") {
			if (v != "auto" && v != "vtu" && v != "vtkhdf" && v != "bezier-vtkhdf")
				throw std::runtime_error("--visualization-format must be auto, vtu, or vtkhdf");
			o.visualization = v == "bezier-vtkhdf" ? "vtkhdf" : v;
		} else if (arg == "--preconditioner") o.preconditioner = Choice(v, arg, {{"amg", TFI_PC_AMG}, {"bjacobi", TFI_PC_BJACOBI}});
		else if (arg == "--precision") o.precision = Choice(v, arg, {{"fp64", TFI_FP64}, {"mixed", TFI_MIXED}});
		else if (arg == "--warm-start")
			o.warm_start = Choice(v, arg, {{"previous", TFI_WARM_PREVIOUS}, {"extrapolate", TFI_WARM_EXTRAPOLATE},
				{"extrapolate2", TFI_WARM_EXTRAPOLATE2}, {"extrapolate3", TFI_WARM_EXTRAPOLATE3},
				{"extrapolate4", TFI_WARM_EXTRAPOLATE4}, {"mpe", TFI_WARM_MPE},
				{"learned", TFI_WARM_LEARNED}});
		else if (arg == "--tangent") o.tangent = Choice(v, arg, {{"exact", TFI_TANGENT_EXACT}, {"legacy", TFI_TANGENT_LEGACY}});
		else if (arg == "--amg-smoother")
			o.smoother = Choice(v, arg, {{"auto", TFI_SMOOTH_AUTO}, {"gauss-seidel", TFI_SMOOTH_GAUSS_SEIDEL},
				{"jacobi", TFI_SMOOTH_JACOBI}, {"chebyshev", TFI_SMOOTH_CHEBYSHEV}});
		else if (arg == "--linear-rtol") o.linear_rtol = PositiveDouble(v, arg);
		else if (arg == "--progress") o.progress = Choice(v, arg, {{"on", 1}, {"off", 0}});
		else if (arg == "--amg-chebyshev-degree") o.chebyshev_degree = PositiveInt(v, arg);
		else if (arg == "--amg-half-smoother") o.half_smoother = Choice(v, arg, {{"on", 1}, {"off", 0}});
		else if (arg == "--restart-length") o.restart_length = PositiveInt(v, arg);
		else if (arg == "--linear-max-iterations") o.linear_maximum_iterations = PositiveInt(v, arg);
		else if (arg == "--chord-polish") o.chord_polish = Choice(v, arg, {{"on", 1}, {"off", 0}});
		else if (arg == "--train-steps") o.train_steps = PositiveInt(v, arg);
		else throw std::runtime_error("unknown option: " + arg);
	}
	if (o.output_every > 0 && o.output.empty()) throw std::runtime_error("--output-every requires --output or legacy OUTPUT");
	if (o.checkpoint_every > 0 && o.checkpoint.empty()) throw std::runtime_error("--checkpoint-every requires --checkpoint");
	return o;
}

void ApplyWallTrace(ResolvedFlowBoundaries& b, const std::vector<std::int32_t>& trace)
{
	for (const auto n : trace) {
		if (!b.velocity_mask.at(n)) ++b.velocity_nodes;
		b.velocity_mask[n] = 1;
		for (int k = 0; k < 3; ++k) b.velocity[3*n + k] = 0.0;
	}
}

double Norm(const std::vector<double>& v, int first, int stride)
{
	double s = 0.0;
	for (std::size_t i = first; i < v.size(); i += stride) s += v[i]*v[i];
	return std::sqrt(s);
}

// The Bezier VTKHDF visualisation (iga_cuda.cu:52-70, 366-386;
// ConfiguredTransportHost.inc:98-140): the merged Bezier mesh and its report
// are built once, then every visualisation time appends the point arrays.
class BezierOutput {
public:
	void Initialize(const FlatDatabase& db, const fs::path& output, bool resume)
	{
#ifdef TFI_HAVE_HDF5
		mesh_ = std::make_unique<bezier::VisualMesh>(BuildBezierVisualMesh(db));
		const auto report = BezierReportPath(output);
		WriteBezierReport(report, mesh_->checks);
		RequireValidBezierMesh(mesh_->checks);
		series_ = std::make_unique<BezierVtkHdfSeries>(VtkHdfOutputPath(output), *mesh_, resume);
		std::cout << "bezier_geometry_points=" << mesh_->points.size() << " local_point_references="
			<< mesh_->checks.local_points << " geometry_report=" << report.string() << " vtkhdf="
			<< series_->path().string() << '\n';
#else
		(void)db; (void)output; (void)resume;
		throw std::runtime_error("--visualization-format vtkhdf needs the HDF5 library, which iga_hip was compiled "
			"without; use --visualization-format vtu");
#endif
	}
	// the flow state [4N] as "velocity" and "pressure"
	void Append(double time, const std::vector<double>& state)
	{
		const std::size_t n = state.size()/4;
		std::vector<double> v(3*n), p(n);
		for (std::size_t i = 0; i < n; ++i) {
			for (int k = 0; k < 3; ++k) v[3*i + k] = state[4*i + k];
			p[i] = state[4*i + 3];
		}
		Append(time, std::vector<PointArray>{{"velocity", 3, std::move(v)}, {"pressure", 1, std::move(p)}});
	}
	void Append(double time, const std::vector<PointArray>& arrays)
	{
#ifdef TFI_HAVE_HDF5
		std::vector<PointField> fields;
		for (const auto& a : arrays) fields.push_back({a.name, a.components, a.values});
		series_->Append(time, fields);
#else
		(void)time; (void)arrays;
#endif
	}
	void Close()
	{
#ifdef TFI_HAVE_HDF5
		if (series_) series_->Close();
#endif
	}

private:
#ifdef TFI_HAVE_HDF5
	std::unique_ptr<bezier::VisualMesh> mesh_;
	std::unique_ptr<BezierVtkHdfSeries> series_;
#endif
};

int NavierStokes(Ctx& c, int argc, char** argv)
{
	const Options o = ParseOptions(argc, argv);
	const auto total_start = Clock::now();
	const FlatDatabase db = ReadDatabase(o.database.string());
	const auto labeled = ReadLabeledHexMesh((o.case_dir/"controlmesh.vtk").string(), db.nodes,
		static_cast<std::uint64_t>(db.elements()));
	const auto wall_trace = WallTraceBasis(db, labeled);
	const auto profile = ReadVelocity((o.case_dir/"initial_velocityfield.txt").string(), db.nodes);
	// Without simulation_config.json the case is the legacy input
	// (iga_cuda.cu:791-795): simulation_parameter.txt and an optional
	// case_config.json, a steady flow with viscosity 0.1 and density 1, node
	// labels as Dirichlet data (wall and inlet velocity, outlet pressure).
	const bool legacy = !fs::exists(o.case_dir/"simulation_config.json");
	SimulationConfig config;
	FlowSystem legacy_system;
	ResolvedFlowBoundaries legacy_bc;
	std::string boundary_config = "simulation_config.json";
	if (legacy) {
		const TransportParameters params = ReadTransportParameters((o.case_dir/"simulation_parameter.txt").string());
		const CaseRules rules = ReadCaseRules((o.case_dir/"case_config.json").string());
		const TransportBoundaries tb = ResolveTransportBoundaries(rules, labeled.labels, params, &profile);
		legacy_bc.velocity_mask = tb.velocity_mask;
		legacy_bc.pressure_mask = tb.pressure_mask;
		legacy_bc.velocity = tb.velocity;
		legacy_bc.pressure = tb.pressure;
		legacy_bc.velocity_nodes = tb.velocity_nodes;
		legacy_bc.pressure_nodes = tb.pressure_nodes;
		legacy_system.viscosity = 0.1;
		legacy_system.density = 1.0;
		legacy_system.time_integration = "steady";
		boundary_config = rules.present ? "case_config.json" : "legacy-defaults";
	} else {
		config = ReadSimulationConfig((o.case_dir/"simulation_config.json").string());
	}
	const FlowSystem& sys = legacy ? legacy_system : FindFlowSystem(config, o.system);
	const bool transient = sys.time_integration == "backward_euler";
	const double dt = transient ? config.dt : 0.0;
	const int physical_steps = transient ? config.steps : 1;
	auto step_config = [&](double time) { return transient ? MaterializeWaveforms(config, o.case_dir, time) : config; };
	OutletBank outlets = legacy ? OutletBank() : OutletBank::FromConfig(config, sys);
	auto outlet_p = outlets.Tractions();
	auto resolve = [&](const SimulationConfig& cfg) {
		ResolvedFlowBoundaries r = legacy ? legacy_bc
			: ResolveFlowBoundaries(cfg, sys, labeled.labels, profile, &outlet_p);
		ApplyWallTrace(r, wall_trace);
		return r;
	};
	ResolvedFlowBoundaries bc = resolve(step_config(0.0));
	if (o.stop_after_step > physical_steps) throw std::runtime_error("--stop-after-step exceeds configured physical steps");
	const int run_end = o.stop_after_step > 0 ? o.stop_after_step : physical_steps;
	std::cout << "boundary_config=" << boundary_config << " viscosity=" << sys.viscosity << " density=" << sys.density
		<< " time_integration=" << (transient ? "backward_euler" : "steady") << " dt=" << dt
		<< " steps=" << physical_steps << " run_end_step=" << run_end << " velocity_nodes=" << bc.velocity_nodes
		<< " pressure_nodes=" << bc.pressure_nodes << '\n';
	if (!transient && (!o.restart.empty() || !o.checkpoint.empty() || o.output_every > 0 || o.checkpoint_every > 0))
		throw std::runtime_error("restart, checkpoint, and time-indexed output require transient flow");

	Mesh m;
	CreateMesh(c, db, m);
	tfi_mesh_stats st{};
	c.Check(tfi_mesh_stats_get(m.mesh, &st), "mesh statistics");
	PrintMesh(db, &st);
	if (st.bad_samples)
		throw std::runtime_error("mesh has " + std::to_string(st.bad_samples)
			+ " non-positive quadrature Jacobians; minimum detJ " + std::to_string(st.minimum_determinant));
	const auto preprocess_end = Clock::now();

	// labelled faces (mass balance) and pressure-traction faces, element-major
	std::vector<std::int32_t> face_el, face_id, face_lab;
	for (int e = 0; e < db.elements(); ++e)
		for (int f = 0; f < 6; ++f)
			if (db.labels[6*e + f] >= 0) {
				face_el.push_back(e);
				face_id.push_back(f);
				face_lab.push_back(db.labels[6*e + f]);
			}
	for (const auto& t : bc.tractions)
		if (std::find(face_lab.begin(), face_lab.end(), t.first) == face_lab.end())
			throw std::runtime_error("pressure traction label " + std::to_string(t.first)
				+ " has no boundary faces in the .ntiga database; repack with iga_pack");

	tfi_flow_options opt;
	tfi_flow_default_options(&opt);
	opt.density = sys.density;
	opt.viscosity = sys.viscosity;
	opt.dt = dt;
	opt.maximum_newton = o.maximum_newton;
	opt.nonlinear_rtol = o.nonlinear_rtol;
	opt.nonlinear_atol = o.nonlinear_atol;
	opt.mass_rtol = o.mass_rtol;
	opt.linear_rtol = o.linear_rtol;
	opt.linear_maximum_iterations = o.linear_maximum_iterations;
	opt.restart = o.restart_length;
	opt.preconditioner = o.preconditioner;
	opt.precision = o.precision;
	opt.warm_start = transient ? o.warm_start : TFI_WARM_PREVIOUS;
	opt.amg_smoother = o.smoother;
	opt.tangent = o.tangent;
	opt.chord_polish = o.chord_polish;
	opt.verbose = o.progress;
	opt.amg_chebyshev_degree = o.chebyshev_degree;
	opt.amg_half_smoother = o.half_smoother;
	Flow flow;
	c.Check(tfi_flow_create(m.mesh, &opt, &flow.flow), "flow solver");
	if (opt.warm_start == TFI_WARM_LEARNED && o.train_steps > 0) {
		tfi_warm_options w{128, 27, o.train_steps, 1e-3, 2, TFI_WARM_MPE};
		c.Check(tfi_flow_warm_configure(flow.flow, &w), "warm start");
	}
	// Inflow phase for the learned predictor: sin / cos of 2 pi t / T and
	// 4 pi t / T, T the period of the first periodic temporal function.
	double period = 0.0;
	for (const auto& fn : config.functions)
		if (fn.period > 0.0) {
			period = fn.period;
			break;
		}
	c.Check(tfi_flow_set_boundary_faces(flow.flow, static_cast<int>(face_el.size()), face_el.data(), face_id.data(),
		face_lab.data()), "boundary faces");

	const std::size_t N = db.nodes;
	std::vector<double> state(4*N, 0.0);
	int start_step = 0;
	if (transient && !o.restart.empty()) {
		const auto meta = ReadCheckpointMeta(o.restart);
		ValidateCheckpoint(meta, db.nodes, physical_steps, dt, sys.density, sys.viscosity);
		state = ReadCheckpointState(o.restart, meta);
		start_step = meta.completed_step;
		outlets.Restore(meta.outlets);
		std::cout << "restart=" << o.restart.string() << " completed_step=" << start_step
			<< " physical_time=" << meta.physical_time << '\n';
	} else if (transient) {
		for (std::size_t n = 0; n < N; ++n) {
			if (bc.velocity_mask[n])
				for (int k = 0; k < 3; ++k) state[4*n + k] = bc.velocity[3*n + k];
			if (bc.pressure_mask[n]) state[4*n + 3] = bc.pressure[n];
		}
	}
	c.Check(tfi_flow_set_state(flow.flow, state.data(), nullptr), "initial state");

	const fs::path mesh_vtk = o.case_dir/"controlmesh.vtk";
	std::vector<std::pair<double, fs::path>> vtu_snaps, text_snaps;
	// visualisation: VTU, or by default for transient runs the Bezier VTKHDF
	// file (ResolveVisualizationFormat, include/VtkOutput.hpp:27-32)
	const bool vtkhdf = o.visualization == "vtkhdf" || (o.visualization == "auto" && transient);
	BezierOutput bezier_out;
	if (!o.output.empty() && vtkhdf) bezier_out.Initialize(db, o.output, !o.restart.empty());
	auto visualise = [&](const fs::path& vtu, double time) {
		if (vtkhdf) {
			bezier_out.Append(time, state);
		} else {
			WriteStateVtu(mesh_vtk, vtu, state, time);
			vtu_snaps.emplace_back(time, vtu);
		}
	};
	if (transient && o.output_every > 0) {
		const auto text = TimeIndexedPath(o.output, start_step);
		WriteStateText(text, state);
		visualise(VtuStepPath(o.output, start_step), start_step*dt);
		text_snaps.emplace_back(start_step*dt, text);
	}
	long long total_linear_iterations = 0;
	double total_assembly = 0.0, total_linear = 0.0;
	std::vector<double> traction(4*N, 0.0);
	const std::vector<std::int32_t> outlet_labels(outlets.labels().begin(), outlets.labels().end());
	for (int step = start_step; step < run_end; ++step) {
		const double time = transient ? (step + 1)*dt : 0.0;
		const SimulationConfig cfg = step_config(time);
		if (period > 0.0) {
			const double w = 2.0*M_PI*time/period;
			const double phase[4] = {std::sin(w), std::cos(w), std::sin(2.0*w), std::cos(2.0*w)};
			c.Check(tfi_flow_begin_step(flow.flow, phase, 4), "begin step");
		} else {
			c.Check(tfi_flow_begin_step(flow.flow, nullptr, 0), "begin step");
		}
		// Outlet fixed point (iga_cuda.cu:927-1084): each pass solves the step
		// with the outlet models' pressures as tractions, from where the last
		// pass ended; the models then turn the outlet flows into new pressures.
		outlets.StartStep();
		const int passes = outlets.empty() ? 1 : 12;
		bool outlets_converged = false;
		tfi_newton_report rep{}, first_rep{};  // last and first outlet pass of the step
		for (int coupling = 0; coupling < passes; ++coupling) {
			outlet_p = outlets.Tractions();
			bc = resolve(cfg);
			std::vector<std::int32_t> te, tf;
			std::vector<double> tp;
			for (std::size_t k = 0; k < face_el.size(); ++k) {
				const auto it = bc.tractions.find(face_lab[k]);
				if (it == bc.tractions.end()) continue;
				te.push_back(face_el[k]);
				tf.push_back(face_id[k]);
				tp.push_back(it->second);
			}
			c.Check(tfi_pressure_traction(m.mesh, static_cast<int>(te.size()), te.data(), tf.data(), tp.data(),
				traction.data()), "pressure traction");
			const bool any_traction = std::any_of(traction.begin(), traction.end(), [](double v) { return v != 0.0; });
			c.Check(tfi_flow_set_boundary(flow.flow, bc.velocity_mask.data(), bc.pressure_mask.data(), bc.velocity.data(),
				bc.pressure.data(), any_traction ? traction.data() : nullptr), "boundary data");
			rep = tfi_newton_report{};
			const int status = tfi_flow_newton(flow.flow, &rep);
			if (coupling == 0) first_rep = rep;
			const double scale = std::sqrt(static_cast<double>(4*N));
			for (int r = 0; r < rep.records; ++r) {
				const auto& rec = rep.record[r];
				const bool last_converged = status == TFI_OK && r == rep.records - 1;
				if (last_converged) {
					std::cout << "step=" << step + 1 << " time=" << time << " converged newton=" << r
						<< " residual_l2=" << rec.residual_l2 << " residual_rms=" << rec.residual_l2/scale
						<< " tolerance=" << rep.tolerance << " absolute_rms_tolerance=" << o.nonlinear_atol
						<< " continuity_l2=" << rec.continuity_l2 << " continuity_sum=" << rec.continuity_sum
						<< " net_boundary_flow=" << rec.net_boundary_flow
						<< " relative_mass_imbalance=" << rec.relative_mass_imbalance
						<< " mass_tolerance=" << o.mass_rtol << '\n';
				} else {
					std::cout << "step=" << step + 1 << " time=" << time << " newton=" << r
						<< " residual_l2=" << rec.residual_l2 << " residual_rms=" << rec.residual_l2/scale
						<< " continuity_l2=" << rec.continuity_l2 << " continuity_sum=" << rec.continuity_sum
						<< " net_boundary_flow=" << rec.net_boundary_flow
						<< " relative_mass_imbalance=" << rec.relative_mass_imbalance << " update_l2=" << rec.update_l2
						<< " linear_iterations=" << rec.linear_iterations << " linear_residual=" << rec.linear_residual
						<< " singular_diagonal_blocks=" << rec.singular_diagonal_blocks << " assembly_s=" << rec.assembly_s
						<< " linear_s=" << rec.linear_s << " step=" << rec.step << '\n';
				}
			}
			if (status != TFI_OK)
				throw std::runtime_error(std::string(tfi_last_error(c.ctx)) + " at physical step " + std::to_string(step + 1));
			total_assembly += rep.assembly_s;
			total_linear += rep.linear_s;
			total_linear_iterations += rep.linear_iterations;
			if (outlets.empty()) {
				outlets_converged = true;
				break;
			}
			std::vector<double> flows(outlets.size());
			c.Check(tfi_flow_boundary_flows(flow.flow, static_cast<int>(outlet_labels.size()), outlet_labels.data(),
				flows.data()), "outlet flows");
			const double change = outlets.Respond(flows, dt);
			std::cout << "step=" << step + 1 << " time=" << time << " outlet_iteration=" << coupling
				<< " pressure_change=" << change << " tolerance=" << outlets.Tolerance() << '\n';
			if (change <= outlets.Tolerance()) {
				outlets.Commit();
				for (std::size_t i = 0; i < outlets.size(); ++i)
					std::cout << "outlet label=" << outlets.labels()[i] << " flow=" << outlets.flow(i)
						<< " pressure=" << outlets.applied(i) << " capacitor_pressure=" << outlets.capacitor(i) << '\n';
				outlets_converged = true;
				break;
			}
			outlets.Relax();
		}
		if (!outlets_converged)
			throw std::runtime_error("outlet fixed-point iteration did not converge at physical step "
				+ std::to_string(step + 1));
		// the prediction is made for the step's first outlet pass only
		if (opt.warm_start != TFI_WARM_PREVIOUS)
			std::cout << "warm_start step=" << step + 1 << " used=" << first_rep.warm_start_used
				<< " initial_ratio=" << first_rep.warm_start_initial_ratio << '\n';
		c.Check(tfi_flow_commit_step(flow.flow), "commit step");
		const int completed = step + 1;
		if (o.output_every > 0 && completed % o.output_every == 0) {
			c.Check(tfi_flow_get_state(flow.flow, state.data()), "state download");
			const auto text = TimeIndexedPath(o.output, completed);
			WriteStateText(text, state);
			visualise(VtuStepPath(o.output, completed), completed*dt);
			text_snaps.emplace_back(completed*dt, text);
		}
		if (!o.checkpoint.empty() && (completed == run_end || (o.checkpoint_every > 0 && completed % o.checkpoint_every == 0))) {
			c.Check(tfi_flow_get_state(flow.flow, state.data()), "state download");
			CheckpointMeta meta;
			meta.nodes = db.nodes;
			meta.completed_step = completed;
			meta.physical_time = completed*dt;
			meta.dt = dt;
			meta.density = sys.density;
			meta.viscosity = sys.viscosity;
			meta.outlets = outlets.Checkpoint();
			WriteCheckpoint(o.checkpoint, meta, state);
			std::cout << "checkpoint=" << o.checkpoint.string() << " completed_step=" << completed << '\n';
		}
	}
	c.Check(tfi_flow_get_state(flow.flow, state.data()), "state download");
	if (!o.output.empty()) {
		WriteStateText(o.output, state);
		if (!vtkhdf) WriteStateVtk(mesh_vtk, o.output.string() + ".vtk", state);
		const double final_time = transient ? run_end*dt : 0.0;
		if (vtkhdf) {
			bezier_out.Append(final_time, state);
		} else {
			WriteStateVtu(mesh_vtk, VtuFinalPath(o.output), state, final_time);
			if (vtu_snaps.empty()) vtu_snaps.emplace_back(final_time, VtuFinalPath(o.output));
			WritePvd(PvdPath(o.output), vtu_snaps);
		}
		if (!text_snaps.empty()) WriteSeries(SeriesPath(o.output), text_snaps);
		bezier_out.Close();
	}
	long long peak = 0, live = 0;
	tfi_allocation_stats(c.ctx, &peak, &live);
	// record name and keys of the reference driver (parsed by its scripts)
	std::cout << "navier_stokes_cuda nodes=" << db.nodes << " elements=" << db.elements() << " steps=" << physical_steps
		<< " run_end_step=" << run_end << " preprocess_s=" << Seconds(total_start, preprocess_end)
		<< " assembly_s=" << total_assembly << " linear_s=" << total_linear
		<< " total_linear_iterations=" << total_linear_iterations << " state_l2=" << Norm(state, 0, 1)
		<< " velocity_l2=" << std::sqrt(std::pow(Norm(state, 0, 4), 2) + std::pow(Norm(state, 1, 4), 2)
			+ std::pow(Norm(state, 2, 4), 2))
		<< " pressure_l2=" << Norm(state, 3, 4) << " gpu_used_gib=" << Gib(static_cast<double>(peak))
		<< " backend=hip" << '\n';
	return 0;
}

void PrintAllocations(std::ostream& out, tfi_ctx* ctx)
{
	long long peak = 0, live = 0;
	if (ctx) tfi_allocation_stats(ctx, &peak, &live);
	out << "cuda_allocations scope=project_device_buffers requested_peak_bytes=" << peak
		<< " requested_live_bytes=" << live << '\n';
}

}  // namespace

// case-check: the host half of navier-stokes without a GPU -- reads the
// database, control mesh, inflow profile and configuration, resolves the
// boundary data at t = dt (or steady) and prints a summary.  No reference
// counterpart; it lets the input path be tested on machines without a GPU.
int CaseCheck(int argc, char** argv)
{
	if (argc < 4 || argc > 6) throw std::runtime_error("usage: iga_hip case-check DATABASE.ntiga CASE_DIR [--system NAME]");
	std::string system;
	if (argc == 6) {
		if (std::string(argv[4]) != "--system") throw std::runtime_error("unknown option: " + std::string(argv[4]));
		system = argv[5];
	} else if (argc == 5) {
		throw std::runtime_error("usage: iga_hip case-check DATABASE.ntiga CASE_DIR [--system NAME]");
	}
	const fs::path case_dir(argv[3]);
	const FlatDatabase db = ReadDatabase(argv[2]);
	const auto labeled = ReadLabeledHexMesh((case_dir/"controlmesh.vtk").string(), db.nodes,
		static_cast<std::uint64_t>(db.elements()));
	const auto trace = WallTraceBasis(db, labeled);
	const auto profile = ReadVelocity((case_dir/"initial_velocityfield.txt").string(), db.nodes);
	const SimulationConfig config = ReadSimulationConfig((case_dir/"simulation_config.json").string());
	const FlowSystem& sys = FindFlowSystem(config, system);
	const bool transient = sys.time_integration == "backward_euler";
	const double time = transient ? config.dt : 0.0;
	const OutletBank outlets = OutletBank::FromConfig(config, sys);
	const auto outlet_p = outlets.Tractions();
	ResolvedFlowBoundaries bc = ResolveFlowBoundaries(transient ? MaterializeWaveforms(config, case_dir, time) : config,
		sys, labeled.labels, profile, &outlet_p);
	ApplyWallTrace(bc, trace);
	double vsum = 0.0, psum = 0.0;
	long long vhash = 0, phash = 0;
	for (std::size_t n = 0; n < db.nodes; ++n) {
		if (bc.velocity_mask[n]) vhash += static_cast<long long>(n);
		if (bc.pressure_mask[n]) phash += static_cast<long long>(n);
		for (int k = 0; k < 3; ++k) vsum += bc.velocity[3*n + k];
		psum += bc.pressure[n];
	}
	std::cout << std::setprecision(17) << "case_check nodes=" << db.nodes << " elements=" << db.elements()
		<< " max_basis=" << db.maximum_basis << " version=" << db.version << " wall_trace=" << trace.size()
		<< " velocity_nodes=" << bc.velocity_nodes << " pressure_nodes=" << bc.pressure_nodes
		<< " velocity_mask_index_sum=" << vhash << " pressure_mask_index_sum=" << phash
		<< " velocity_target_sum=" << vsum << " pressure_target_sum=" << psum << " time=" << time
		<< " density=" << sys.density << " viscosity=" << sys.viscosity << " traction_labels=";
	bool first = true;
	for (const auto& t : bc.tractions) {
		std::cout << (first ? "" : ",") << t.first << ':' << t.second;
		first = false;
	}
	std::cout << '\n';
	for (std::size_t i = 0; i < outlets.size(); ++i)
		std::cout << "outlet label=" << outlets.labels()[i] << " kind=" << outlets.kind(i) << " pressure="
			<< outlets.applied(i) << " capacitor_pressure=" << outlets.capacitor(i) << '\n';
	return 0;
}

struct TransportOptions {
	fs::path database, case_dir, output, velocity;
	int steps = -1;
	double linear_rtol = 1e-8;
	int preconditioner = TFI_PC_BJACOBI;
};

// iga_cuda transport DATABASE CASE_DIR [STEPS] [OUTPUT] [VELOCITY] (iga_cuda.cu:623-625),
// plus --linear-rtol R (the FGMRES relative tolerance, 1e-8 as in iga_cuda.cu:697)
TransportOptions ParseTransportOptions(int argc, char** argv)
{
	const char* usage = "usage: iga_hip transport DATABASE.ntiga CASE_DIR [STEPS] [OUTPUT] [VELOCITY] [--linear-rtol R] "
		"[--preconditioner amg|bjacobi]";
	std::vector<std::string> pos;
	TransportOptions o;
	for (int i = 2; i < argc; ++i) {
		const std::string arg(argv[i]);
		if (arg == "--linear-rtol") {
			if (i + 1 >= argc) throw std::runtime_error(usage);
			char* end = nullptr;
			o.linear_rtol = std::strtod(argv[++i], &end);
			if (!end || *end || !(o.linear_rtol > 0.0)) throw std::runtime_error("--linear-rtol requires a positive number");
		} else if (arg == "--preconditioner") {
			if (i + 1 >= argc) throw std::runtime_error(usage);
			const std::string v(argv[++i]);
			if (v == "amg") o.preconditioner = TFI_PC_AMG;
			else if (v == "bjacobi") o.preconditioner = TFI_PC_BJACOBI;
			else throw std::runtime_error("--preconditioner must be amg|bjacobi");
		} else if (arg.rfind("--", 0) == 0) {
			throw std::runtime_error("unknown option: " + arg);
		} else {
			pos.push_back(arg);
		}
	}
	if (pos.size() < 2 || pos.size() > 5) throw std::runtime_error(usage);
	o.database = pos[0];
	o.case_dir = pos[1];
	if (pos.size() >= 3) {
		char* end = nullptr;
		const long v = std::strtol(pos[2].c_str(), &end, 10);
		if (!end || *end || v < 0) throw std::runtime_error("STEPS must be a non-negative integer");
		o.steps = static_cast<int>(v);
	}
	if (pos.size() >= 4) o.output = pos[3];
	o.velocity = pos.size() >= 5 ? fs::path(pos[4]) : o.case_dir/"initial_velocityfield.txt";
	return o;
}

// Inputs of the transport command, read (and validated) before the device is touched.
struct TransportCase {
	TransportOptions o;
	Clock::time_point start;
	FlatDatabase db;
	TransportParameters p;
	std::vector<std::array<double, 3>> velocity;
	CaseRules rules;
	TransportBoundaries bc;
};

TransportCase LoadTransportCase(int argc, char** argv)
{
	TransportCase t;
	t.o = ParseTransportOptions(argc, argv);
	t.start = Clock::now();
	t.db = ReadDatabase(t.o.database.string());
	t.p = ReadTransportParameters((t.o.case_dir/"simulation_parameter.txt").string());
	if (t.o.steps >= 0) t.p.steps = t.o.steps;
	const auto labeled = ReadLabeledHexMesh((t.o.case_dir/"controlmesh.vtk").string(), t.db.nodes,
		static_cast<std::uint64_t>(t.db.elements()));
	t.velocity = ReadVelocity(t.o.velocity.string(), t.db.nodes);
	t.rules = ReadCaseRules((t.o.case_dir/"case_config.json").string());
	t.bc = ResolveTransportBoundaries(t.rules, labeled.labels, t.p);
	return t;
}

// transport-check DATABASE CASE_DIR [STEPS] [OUTPUT] [VELOCITY]: the transport
// inputs resolved without a GPU (no reference counterpart).
int TransportCheck(int argc, char** argv)
{
	std::vector<char*> args(argv, argv + argc);
	const TransportCase t = LoadTransportCase(argc, args.data());
	long long index_sum = 0;
	double n0 = 0.0, nplus = 0.0;
	for (std::size_t i = 0; i < t.bc.constrained.size(); ++i) {
		if (t.bc.constrained[i]) index_sum += static_cast<long long>(i);
		n0 += t.bc.n0[i];
		nplus += t.bc.nplus[i];
	}
	const TransportParameters& p = t.p;
	std::cout << std::setprecision(17) << "transport_check nodes=" << t.db.nodes << " elements=" << t.db.elements()
		<< " boundary_config=" << (t.rules.present ? "case_config.json" : "legacy-defaults")
		<< " transport_nodes=" << t.bc.transport_nodes << " velocity_nodes=" << t.bc.velocity_nodes
		<< " pressure_nodes=" << t.bc.pressure_nodes << " constrained_index_sum=" << index_sum << " n0_sum=" << n0
		<< " nplus_sum=" << nplus << " D=" << p.diffusion << " vplus=" << p.vplus << " vminus=" << p.vminus
		<< " kplus=" << p.kplus << " kminus=" << p.kminus << " detach_plus=" << p.detach_plus
		<< " detach_minus=" << p.detach_minus << " dt=" << p.dt << " nstep=" << p.steps << " N0bc=" << p.n0_bc
		<< " Nplusbc=" << p.nplus_bc << " Nminusbc=" << p.nminus_bc << " artificial_diffusion=" << p.artificial_diffusion
		<< '\n';
	return 0;
}

int Transport(Ctx& c, const TransportCase& t)
{
	const TransportOptions& o = t.o;
	const auto total_start = t.start;
	const FlatDatabase& db = t.db;
	const TransportParameters& p = t.p;
	const auto& velocity = t.velocity;
	const TransportBoundaries& bc = t.bc;
	std::cout << "boundary_config=" << (t.rules.present ? "case_config.json" : "legacy-defaults")
		<< " transport_nodes=" << bc.transport_nodes << '\n';
	Mesh m;
	CreateMesh(c, db, m);
	tfi_mesh_stats st{};
	c.Check(tfi_mesh_stats_get(m.mesh, &st), "mesh statistics");
	PrintMesh(db, &st);
	if (st.bad_samples)
		throw std::runtime_error("mesh has " + std::to_string(st.bad_samples)
			+ " non-positive quadrature Jacobians; minimum detJ " + std::to_string(st.minimum_determinant));
	const auto preprocess_end = Clock::now();
	std::vector<double> vel(3*db.nodes);
	for (std::size_t n = 0; n < db.nodes; ++n)
		for (int d = 0; d < 3; ++d) vel[3*n + d] = velocity[n][d];
	const tfi_transport_params tp{p.diffusion, p.vplus, p.vminus, p.kplus, p.kminus, p.detach_plus, p.detach_minus,
		p.dt, p.artificial_diffusion};
	struct Handle {
		tfi_transport* t = nullptr;
		~Handle() { tfi_transport_destroy(t); }
	} tr;
	const auto assembly_start = Clock::now();
	c.Check(tfi_transport_create(m.mesh, &tp, vel.data(), bc.constrained.data(), &tr.t), "transport assembly");
	c.Check(tfi_transport_set_preconditioner(tr.t, o.preconditioner), "transport preconditioner");
	c.Check(tfi_transport_set_boundary(tr.t, bc.n0.data(), bc.nplus.data()), "transport boundary values");
	c.Check(tfi_transport_apply_boundary(tr.t), "transport boundary state");
	const auto assembly_end = Clock::now();
	long long total_iterations = 0;
	int singular = 0;
	c.Check(tfi_transport_matrices(tr.t, nullptr, nullptr, &singular), "transport block inverses");
	const auto solve_start = Clock::now();
	for (int step = 0; step < p.steps; ++step) {
		int it = 0;
		double res = 0.0;
		if (tfi_transport_step(tr.t, step == 0 ? 1 : 0, 10000, o.linear_rtol, &it, &res) != TFI_OK)
			throw std::runtime_error(std::string(tfi_last_error(c.ctx)) + " at step " + std::to_string(step));
		total_iterations += it;
		if ((step + 1) % 25 == 0 || step + 1 == p.steps)
			std::cout << "transport step=" << step + 1 << '/' << p.steps << " iterations=" << it << " residual=" << res
				<< '\n';
	}
	const auto solve_end = Clock::now();
	std::vector<double> state(2*db.nodes);
	c.Check(tfi_transport_get_state(tr.t, state.data()), "transport state download");
	double norm = 0.0;
	for (double v : state) norm += v*v;
	norm = std::sqrt(norm);
	if (!o.output.empty()) {
		WriteTransportText(o.output, state);
		WriteTransportVtk(o.case_dir/"controlmesh.vtk", o.output.string() + ".vtk", state);
	}
	long long peak = 0, live = 0;
	c.Check(tfi_allocation_stats(c.ctx, &peak, &live), "allocation statistics");
	std::cout << "transport_cuda nodes=" << db.nodes << " elements=" << db.elements() << " steps=" << p.steps
		<< " preprocess_s=" << Seconds(total_start, preprocess_end) << " assembly_s=" << Seconds(assembly_start, assembly_end)
		<< " solve_s=" << Seconds(solve_start, solve_end) << " total_iterations=" << total_iterations
		<< " final_l2=" << norm << " singular_diagonal_blocks=" << singular << " gpu_used_gib=" << Gib(peak)
		<< " backend=hip\n";
	return 0;
}

struct SolveOptions {
	fs::path database, case_dir, output, velocity, checkpoint, restart;
	std::string system;
	int output_every = 0, checkpoint_every = 0, stop_after_step = 0;
	double linear_rtol = 1e-8;
	int preconditioner = TFI_PC_BJACOBI;
	std::string visualization = "vtkhdf";  // auto: the Bezier VTKHDF series (runs are transient)
};

// ParseCudaConfiguredTransportOptions (ConfiguredTransportHost.inc:13-48), plus
// --linear-rtol R (the GMRES relative tolerance, 1e-8 as in :391-392)
SolveOptions ParseSolveOptions(int argc, char** argv)
{
	if (argc < 4)
		throw std::runtime_error("usage: iga_hip solve DATABASE.ntiga CASE_DIR [SYSTEM] [OUTPUT] [VELOCITY] "
			"[--system NAME] [--output PATH] [--velocity PATH] [--output-every N] [--checkpoint PREFIX] "
			"[--checkpoint-every N] [--restart PREFIX] [--stop-after-step N] [--visualization-format auto|vtu|vtkhdf] "
			"[--linear-rtol R] [--preconditioner bjacobi|amg]");
	SolveOptions o;
	o.database = argv[2];
	o.case_dir = argv[3];
	int positional = 0;
	for (int i = 4; i < argc; ++i) {
		const std::string arg(argv[i]);
		if (arg.rfind("--", 0) != 0) {
			if (positional == 0) o.system = arg;
			else if (positional == 1) o.output = arg;
			else if (positional == 2) o.velocity = arg;
			else throw std::runtime_error("too many positional arguments");
			++positional;
			continue;
		}
		if (i + 1 >= argc) throw std::runtime_error(arg + " requires a value");
		const std::string v(argv[++i]);
		if (arg == "--system") o.system = v;
		else if (arg == "--output") o.output = v;
		else if (arg == "--velocity") o.velocity = v;
		else if (arg == "--output-every") o.output_every = PositiveInt(v, arg);
		else if (arg == "--checkpoint") o.checkpoint = v;
		else if (arg == "--checkpoint-every") o.checkpoint_every = PositiveInt(v, arg);
		else if (arg == "--restart") o.restart = v;
		else if (arg == "--stop-after-step") o.stop_after_step = PositiveInt(v, arg);
		else if (arg == "--visualization-format") {
			if (v != "auto" && v != "vtu" && v != "vtkhdf" && v != "bezier-vtkhdf")
				throw std::runtime_error("--visualization-format must be auto, vtu, or vtkhdf");
			o.visualization = v == "vtu" ? "vtu" : "vtkhdf";  // auto: backward Euler runs are transient
		} else if (arg == "--linear-rtol") o.linear_rtol = PositiveDouble(v, arg);
		else if (arg == "--preconditioner") o.preconditioner = Choice(v, arg, {{"bjacobi", TFI_PC_BJACOBI}, {"amg", TFI_PC_AMG}});
		else throw std::runtime_error("unknown option: " + arg);
	}
	if (o.output_every > 0 && o.output.empty()) throw std::runtime_error("--output-every requires --output");
	if (o.checkpoint_every > 0 && o.checkpoint.empty()) throw std::runtime_error("--checkpoint-every requires --checkpoint");
	return o;
}

// Inputs of the solve command, read (and validated) before the device is touched.
struct SolveCase {
	SolveOptions o;
	Clock::time_point start;
	FlatDatabase db;
	LinearConfig config;
	CompiledSystem system;
	ScalarBoundaries bc;
	std::vector<double> velocity;  // [3N] prescribed velocity (empty for a snapshot series)
	std::vector<VelocitySnapshot> snapshots;
	VelocitySourceDef source;
	std::vector<double> state;     // [N*F] initial or restart state
	int start_step = 0, end_step = 0;
};

SolveCase LoadSolveCase(int argc, char** argv)
{
	SolveCase t;
	t.o = ParseSolveOptions(argc, argv);
	t.start = Clock::now();
	t.db = ReadDatabase(t.o.database.string());
	t.config = ReadLinearConfig((t.o.case_dir/"simulation_config.json").string());
	t.system = CompileSystem(t.config, t.o.system);
	const CompiledSystem& s = t.system;
	if (s.velocity_source != "prescribed") {
		if (!t.o.velocity.empty())
			throw std::runtime_error("a positional VELOCITY file cannot override a configured snapshot_series");
		const auto it = std::find_if(t.config.sources.begin(), t.config.sources.end(),
			[&](const VelocitySourceDef& d) { return d.name == s.velocity_source; });
		if (it == t.config.sources.end()) throw std::runtime_error("unknown velocity source '" + s.velocity_source + "'");
		t.source = *it;
		t.snapshots = ReadVelocityManifest(t.o.case_dir/t.source.manifest);
	}
	if (t.o.stop_after_step > s.steps) throw std::runtime_error("--stop-after-step exceeds configured transport steps");
	t.end_step = t.o.stop_after_step > 0 ? t.o.stop_after_step : s.steps;
	const auto labeled = ReadLabeledHexMesh((t.o.case_dir/"controlmesh.vtk").string(), t.db.nodes,
		static_cast<std::uint64_t>(t.db.elements()));
	if (t.snapshots.empty()) {
		const auto vel = ReadVelocity(
			(t.o.velocity.empty() ? t.o.case_dir/"initial_velocityfield.txt" : t.o.velocity).string(), t.db.nodes);
		t.velocity.resize(3*t.db.nodes);
		for (std::size_t n = 0; n < t.db.nodes; ++n)
			for (int d = 0; d < 3; ++d) t.velocity[3*n + d] = vel[n][d];
	}
	t.bc = ResolveScalarBoundaries(t.config, s, labeled.labels, t.db);
	const std::size_t F = s.fields.size();
	t.state.resize(t.db.nodes*F);
	for (std::size_t n = 0; n < t.db.nodes; ++n)
		for (std::size_t f = 0; f < F; ++f)
			t.state[n*F + f] = t.bc.constrained[n*F + f] ? t.bc.value[n*F + f] : s.initial[f];
	if (!t.o.restart.empty()) {
		const TransportCheckpointMeta m = ReadTransportCheckpoint(t.o.restart, &t.state);
		ValidateTransportCheckpoint(m, t.db.nodes, s);
		t.start_step = m.completed_step;
	}
	return t;
}

// solve-check DATABASE CASE_DIR [options]: the compiled system and its
// boundaries without a GPU (no reference counterpart)
int SolveCheck(int argc, char** argv)
{
	const SolveCase t = LoadSolveCase(argc, argv);
	const CompiledSystem& s = t.system;
	std::cout << std::setprecision(17) << "solve_check system=" << s.name << " fields=";
	for (std::size_t i = 0; i < s.fields.size(); ++i) std::cout << (i ? "," : "") << s.fields[i];
	std::cout << " velocity_source=" << s.velocity_source << " dt=" << s.dt << " steps=" << s.steps
		<< " terms=" << s.terms.size() << " constrained_dofs=" << t.bc.constrained_dofs
		<< " surfaces=" << t.bc.surface_element.size() << '\n';
	for (const auto& term : s.terms)
		std::cout << "term kind=" << term.kind << " equation=" << term.equation << " trial=" << term.trial
			<< " coefficient=" << term.coefficient << '\n';
	for (std::size_t f = 0; f < s.fields.size(); ++f)
		std::cout << "field index=" << f << " supg=" << s.supg[f] << " tau_diffusivity=" << s.tau_diffusivity[f]
			<< " initial=" << s.initial[f] << '\n';
	long long index_sum = 0;
	double value_sum = 0.0;
	for (std::size_t i = 0; i < t.bc.constrained.size(); ++i)
		if (t.bc.constrained[i]) {
			index_sum += static_cast<long long>(i);
			value_sum += t.bc.value[i];
		}
	std::cout << "boundary constrained_index_sum=" << index_sum << " value_sum=" << value_sum << '\n';
	// derived physiology arrays of an interleaved state in CASE_DIR/physiology_state.txt
	std::ifstream state_file(t.o.case_dir/"physiology_state.txt");
	if (state_file) {
		std::vector<double> values;
		double v = 0.0;
		while (state_file >> v) values.push_back(v);
		for (const auto& a : PhysiologyArrays(t.config.physiology, s.fields, values)) {
			std::cout << "derived name=" << a.name;
			for (double x : a.values) std::cout << ' ' << x;
			std::cout << '\n';
		}
	}
	return 0;
}

// velocity_at (ConfiguredTransportHost.inc:343-357): the snapshot series
// interpolated linearly at `time`, read from the case directory
std::vector<double> SeriesVelocity(const SolveCase& t, double time)
{
	const VelocityInterpolation w = ResolveVelocityInterpolation(t.snapshots, time, t.source.out_of_range);
	const auto lower = ReadVelocity((t.o.case_dir/t.snapshots[w.lower].file).string(), t.db.nodes);
	std::vector<double> out(3*t.db.nodes);
	for (std::size_t n = 0; n < t.db.nodes; ++n)
		for (int d = 0; d < 3; ++d) out[3*n + d] = lower[n][d];
	if (w.lower == w.upper) return out;
	const auto upper = ReadVelocity((t.o.case_dir/t.snapshots[w.upper].file).string(), t.db.nodes);
	for (std::size_t n = 0; n < t.db.nodes; ++n)
		for (int d = 0; d < 3; ++d) out[3*n + d] = (1.0 - w.upper_weight)*out[3*n + d] + w.upper_weight*upper[n][d];
	return out;
}

int Solve(Ctx& c, const SolveCase& t)
{
	const SolveOptions& o = t.o;
	const FlatDatabase& db = t.db;
	const CompiledSystem& s = t.system;
	const int F = static_cast<int>(s.fields.size());
	Mesh m;
	CreateMesh(c, db, m);
	tfi_mesh_stats st{};
	c.Check(tfi_mesh_stats_get(m.mesh, &st), "mesh statistics");
	PrintMesh(db, &st);
	if (st.bad_samples)
		throw std::runtime_error("mesh has " + std::to_string(st.bad_samples)
			+ " non-positive quadrature Jacobians; minimum detJ " + std::to_string(st.minimum_determinant));
	const auto preprocess_end = Clock::now();
	if (!o.restart.empty())
		std::cout << "restart=" << o.restart.string() << " completed_step=" << t.start_step
			<< " physical_time=" << t.start_step*s.dt << '\n';
	tfi_solve_system sys{};
	sys.fields = F;
	sys.term_count = static_cast<int>(s.terms.size());
	sys.terms = s.terms.data();
	sys.supg = s.supg.data();
	sys.tau_diffusivity = s.tau_diffusivity.data();
	sys.dt = s.dt;
	sys.surface_count = static_cast<int>(t.bc.surface_element.size());
	sys.surface_element = t.bc.surface_element.data();
	sys.surface_face = t.bc.surface_face.data();
	sys.surface_flux = t.bc.surface_flux.data();
	sys.surface_robin = t.bc.surface_robin.data();
	sys.surface_robin_source = t.bc.surface_robin_source.data();
	struct Handle {
		tfi_solve* s = nullptr;
		~Handle() { tfi_solve_destroy(s); }
	} h;
	const bool series = !t.snapshots.empty();
	const auto assembly_start = Clock::now();
	// with a snapshot series the operators are assembled before every step
	// for the velocity at its end time; the first assembly uses the first one
	const std::vector<double> first_velocity = series ? SeriesVelocity(t, (t.start_step + 1)*s.dt) : t.velocity;
	c.Check(tfi_solve_create(m.mesh, &sys, first_velocity.data(), t.bc.constrained.data(), &h.s),
		"configured transport assembly");
	c.Check(tfi_solve_set_preconditioner(h.s, o.preconditioner), "configured transport preconditioner");
	c.Check(tfi_solve_set_boundary(h.s, t.bc.value.data()), "configured transport boundary values");
	c.Check(tfi_solve_set_state(h.s, t.state.data()), "configured transport state");
	double assembly_s = Seconds(assembly_start, Clock::now());
	int singular = 0;
	c.Check(tfi_solve_matrices(h.s, nullptr, nullptr, nullptr, &singular), "configured transport block inverses");
	std::vector<double> state(t.state);
	const fs::path mesh_vtk = o.case_dir/"controlmesh.vtk";
	std::vector<std::pair<double, fs::path>> snapshots;
	const bool vtkhdf = o.visualization == "vtkhdf";
	BezierOutput bezier_out;
	if (!o.output.empty() && vtkhdf) bezier_out.Initialize(db, o.output, !o.restart.empty());
	std::function<void(double, const std::vector<PointArray>&)> visual;
	if (vtkhdf) visual = [&](double time, const std::vector<PointArray>& arrays) { bezier_out.Append(time, arrays); };
	auto write_output = [&](int step) {
		c.Check(tfi_solve_get_state(h.s, state.data()), "configured transport state download");
		const fs::path vtu = VtuStepPath(o.output, step);
		WriteConfiguredOutput(TimeIndexedPath(o.output, step), s.fields, state, mesh_vtk, vtu, step*s.dt,
			t.config.physiology, visual);
		if (!vtkhdf) snapshots.push_back({step*s.dt, vtu});
	};
	if (o.output_every > 0) write_output(t.start_step);
	long long total_iterations = 0;
	double linear_s = 0.0;
	const auto solve_start = Clock::now();
	for (int step = t.start_step; step < t.end_step; ++step) {
		int it = 0;
		double res = 0.0;
		if (series && step > t.start_step) {
			const auto as = Clock::now();
			const auto vel = SeriesVelocity(t, (step + 1)*s.dt);
			c.Check(tfi_solve_set_velocity(h.s, vel.data()), "configured transport assembly");
			assembly_s += Seconds(as, Clock::now());
		}
		const auto ls = Clock::now();
		if (tfi_solve_step(h.s, step == 0 ? 1 : 0, 10000, o.linear_rtol, &it, &res) != TFI_OK)
			throw std::runtime_error("configured transport GMRES failed at step " + std::to_string(step) + ": "
				+ tfi_last_error(c.ctx));
		linear_s += Seconds(ls, Clock::now());
		total_iterations += it;
		const int completed = step + 1;
		if (o.output_every > 0 && completed % o.output_every == 0) write_output(completed);
		if (!o.checkpoint.empty()
			&& (completed == t.end_step || (o.checkpoint_every > 0 && completed % o.checkpoint_every == 0))) {
			c.Check(tfi_solve_get_state(h.s, state.data()), "configured transport state download");
			TransportCheckpointMeta meta;
			meta.nodes = db.nodes;
			meta.fields = s.fields;
			meta.system = s.name;
			meta.velocity_source = s.velocity_source;
			meta.completed_step = completed;
			meta.physical_time = completed*s.dt;
			meta.dt = s.dt;
			WriteTransportCheckpoint(o.checkpoint, meta, state);
			std::cout << "checkpoint=" << o.checkpoint.string() << " completed_step=" << completed << '\n';
		}
	}
	const auto solve_end = Clock::now();
	c.Check(tfi_solve_get_state(h.s, state.data()), "configured transport state download");
	double norm = 0.0;
	for (double v : state) norm += v*v;
	norm = std::sqrt(norm);
	if (!o.output.empty()) {
		WriteConfiguredOutput(o.output, s.fields, state, mesh_vtk, VtuFinalPath(o.output), t.end_step*s.dt,
			t.config.physiology, visual);
		if (!vtkhdf) {
			if (snapshots.empty()) snapshots.push_back({t.end_step*s.dt, VtuFinalPath(o.output)});
			WritePvd(PvdPath(o.output), snapshots);
		}
		WritePhysiologyManifest(o.output.parent_path()/"physiology_fields.json", t.config.physiology, s.fields);
		bezier_out.Close();
	}
	long long peak = 0, live = 0;
	c.Check(tfi_allocation_stats(c.ctx, &peak, &live), "allocation statistics");
	std::cout << "iga_cuda_solve system=" << s.name << " fields=" << F << " steps=" << s.steps
		<< " run_end_step=" << t.end_step << " preprocess_s=" << Seconds(t.start, preprocess_end)
		<< " assembly_s=" << assembly_s << " solve_s=" << linear_s
		<< " time_loop_s=" << Seconds(solve_start, solve_end) << " total_iterations=" << total_iterations
		<< " final_l2=" << norm << " singular_diagonal_blocks=" << singular << " gpu_used_gib=" << Gib(peak)
		<< " backend=hip\n";
	return 0;
}

int main(int argc, char** argv)
{
	std::unique_ptr<Ctx> ctx;
	try {
		if (argc < 2)
			throw std::runtime_error("usage: iga_hip device-info|mesh-check|transport|solve|navier-stokes|case-check ...");
		const std::string command(argv[1]);
		if (command == "case-check") return CaseCheck(argc, argv);
		if (command == "transport-check") return TransportCheck(argc, argv);
		if (command == "solve-check") return SolveCheck(argc, argv);
		// usage errors before the device is touched
		if (command == "navier-stokes") ParseOptions(argc, argv);
		std::unique_ptr<TransportCase> transport_case;
		if (command == "transport") transport_case = std::make_unique<TransportCase>(LoadTransportCase(argc, argv));
		std::unique_ptr<SolveCase> solve_case;
		if (command == "solve") solve_case = std::make_unique<SolveCase>(LoadSolveCase(argc, argv));
		ctx = std::make_unique<Ctx>(0);
		PrintDevice(*ctx);
		if (command == "device-info") return 0;
		if (command == "mesh-check") {
			if (argc != 3) throw std::runtime_error("usage: iga_hip mesh-check DATABASE.ntiga");
			return MeshCheck(*ctx, argv[2]);
		}
		int status = 0;
		if (command == "navier-stokes") status = NavierStokes(*ctx, argc, argv);
		else if (command == "transport") status = Transport(*ctx, *transport_case);
		else if (command == "solve") status = Solve(*ctx, *solve_case);
		else throw std::runtime_error("unknown command: " + command);
		PrintAllocations(std::cout, ctx->ctx);
		return status;
	} catch (const std::exception& e) {
		std::cerr << "iga_hip: " << e.what() << '\n';
		PrintAllocations(std::cerr, ctx ? ctx->ctx : nullptr);
		return 1;
	}
}
