// One photon step and one e+- step as device functions: the per-source body
// of the photon front (PhotonFront.inl:137-397: selection, transport, final
// state and materialisation) and of the lepton front (LeptonFront.inl:226-550:
// selection, transport, vertex reselection, final state and materialisation),
// together with their profile contributions (ProfileAccumulation.inl:228-412).
// Children carry their index k among the source's children in history_id; the
// scatter kernel replaces it with first_history + child offset + k.
#pragma once

#include "em/magnetic.hpp"
#include "em/moliere.hpp"
#include "em/physics.hpp"

namespace c8mi::em::phys {

// ---------------------------------------------------------------- photon ----

C8MI_HD void photon_step(EmContext const& c, Accumulator const& acc, Particle const& p,
                Outputs& o) {
  Interaction in;
  Fallback fb{false, 0};
  select_interaction(c, p, in, fb);
  if (fb.flag) {
    record_fallback(acc, p, fb.reason, in.process_id, kStageSelection);
    return;
  }
  // transport_photon (PhotonTransportStep.inl:57-261)
  Particle start = in.particle;
  auto transport_fallback = [&](int reason) {
    record_fallback(acc, in.particle, reason, in.process_id, kStageTransport);
  };
  if (start.pid != 22) return transport_fallback(kFrUnsupportedParticle);
  if (in.status != kSelected && in.status != kNoDiscrete && in.status != kParticleCut)
    return transport_fallback(kFrInvalidFinalState);
  env::Layer const layer = env::query(c, start.position_m, start.direction, start.environment_node);
  if (layer.status != env::kSuccess) return transport_fallback(kFrUnsupportedGeometry);
  start.environment_node = layer.node;
  Particle end = start;
  int limit;
  double cut_deposit = 0.;
  bool observed_before_cut = false;
  bool vertex = false;
  double vertex_density = 0.;
  if (in.status == kParticleCut) {
    end.step_id++;
    limit = kPlCut;
    cut_deposit = start.energy_gev;
  } else {
    env::Boundary const bnd = env::distance_to_boundary(c, start.position_m, start.direction, start.environment_node);
    double const obs = env::observation_distance(c, start.position_m, start.direction);
    bool const has_b = bnd.status == env::kSuccess, has_o = obs > 0.;
    if (!has_b && !has_o) return transport_fallback(kFrUnsupportedGeometry);
    bool const obs_wins = has_o && (!has_b || obs < bnd.distance);
    double const bd = obs_wins ? obs : bnd.distance;
    env::Grammage const bg = env::competition_grammage(c, layer.node, start.position_m, start.direction, bd);
    if (bg.status != env::kSuccess) return transport_fallback(kFrGrammage);
    bool const reaches = in.status == kSelected && isfinite(in.interaction_grammage) &&
                         in.interaction_grammage >= 0. && in.interaction_grammage <= bg.value;
    if (reaches) {
      env::Grammage const dist =
          env::arclength(c, layer.node, start.position_m, start.direction, in.interaction_grammage);
      if (dist.status != env::kSuccess || dist.value > bd * (1. + 1.e-12))
        return transport_fallback(kFrInverseGrammage);
      for (int a = 0; a < 3; ++a) end.position_m[a] = start.position_m[a] + dist.value * start.direction[a];
      end.time_s = start.time_s + dist.value / kC;
      env::Layer const v = env::query(c, end.position_m, end.direction, start.environment_node);
      if (v.status != env::kSuccess || !isfinite(v.density) || !(v.density > 0.))
        return transport_fallback(kFrVertexLookup);
      vertex_density = v.density;
      limit = kPlInteraction;
      if (end.time_s > kCutTime) {
        end.step_id++;
        limit = kPlCut;
        cut_deposit = start.energy_gev;
      } else {
        vertex = true;
      }
    } else {
      if (!isfinite(bg.value)) return transport_fallback(kFrGrammage);
      for (int a = 0; a < 3; ++a) end.position_m[a] = start.position_m[a] + bd * start.direction[a];
      end.time_s = start.time_s + bd / kC;
      end.step_id++;
      if (obs_wins) {
        limit = kPlObservation;
      } else {
        std::int32_t const node = bnd.next_node;
        end.environment_node = node;
        c8mi_env_model const* m = env::model_of(c, node);
        if (!m) {
          limit = kPlEscaped;
        } else {
          env::Layer const next = env::query(c, end.position_m, end.direction, node);
          if (next.status != env::kSuccess || node == layer.node)
            return transport_fallback(kFrUnsupportedGeometry);
          limit = kPlLayer;
          end.medium_id = m->medium_id;
        }
      }
      if (end.time_s > kCutTime) {
        observed_before_cut = limit == kPlObservation;
        limit = kPlCut;
        cut_deposit = start.energy_gev;
      }
    }
  }
  o.step = true;
  // accumulate_photon_profile_step
  double const w = start.weight;
  double const gs = project_grammage(acc, start.position_m);
  double const ge = project_grammage(acc, end.position_m);
  particle_profile(acc, start.pid, gs, ge, w);
  point_energy_profile(acc, ge, cut_deposit * w);
  count(acc, kSteps);
  if (limit == kPlCut) {
    count(acc, kPhotonCuts);
    ledger(acc, kCutDeposit, cut_deposit * w);
    if (observed_before_cut) {
      ledger(acc, kOverlap, cut_deposit * w);
      terminal_energy(acc, end.energy_gev * w, true);
      record_observation(acc, end);
    }
    return;
  }
  if (limit == kPlObservation || limit == kPlEscaped) {
    terminal_energy(acc, end.energy_gev * w, limit == kPlObservation);
    record_observation(acc, end);
    return;
  }
  if (limit == kPlLayer) {
    emit(o, end, kSlotNextContinue);
    return;
  }
  (void)vertex;
  // vertex: classify_photon_final_state (PhotonFinalStateStep.inl:273-558)
  o.interaction = true;
  Particle const& parent = end;
  auto final_fallback = [&](int reason) {
    record_fallback(acc, parent, reason, in.process_id, kStageFinal);
  };
  if (in.deferred_reason >= 0)
    return final_fallback(requires_selected_loss(in.deferred_reason) ? in.deferred_reason
                                                                     : kFrInvalidFinalState);
  int const cap = capability(parent.pid, in.process_id);
  if (cap != kCapPair && cap != kCapMuPair && cap != kCapCompton && cap != kCapPhotoelectric)
    return final_fallback(cap == kCapCpuOnly ? kFrCpuOnly : kFrNotImplemented);
  if (cap == kCapMuPair) {
    // muon pair (PhotonFinalStateStep.inl:326-382, 679-731): the split x = E(mu+)/E
    // from the native CDF, polar angle m/E for both muons, opposite azimuths; no
    // thinning (the muons are not EM) and no LPM veto
    double const mu_mass = c.tables.constants.muon_mass_mev / 1000.;
    if (!isfinite(parent.energy_gev) || !isfinite(parent.weight) || parent.weight < 0. || !(mu_mass > 0.) ||
        parent.generation == 0xffffffffU || parent.step_id == 0xffffffffffffffffULL)
      return final_fallback(kFrInvalidFinalState);
    double const su = uniform(c, parent, kPhotonMuPair, kMuPairSplitDraw);
    Tq const split = dq_loss_fraction(c, 22, kPhotonMuPairFinalState, in.component_hash,
                                      parent.energy_gev * 1000., su);
    if (split.status != kTsSuccess) return final_fallback(kFrInvalidFinalState);
    double const minus_total = (1. - split.value) * parent.energy_gev;
    double const plus_total = split.value * parent.energy_gev;
    if (!isfinite(split.value) || !(minus_total > 0.) || !(plus_total > 0.))
      return final_fallback(kFrInvalidFinalState);
    double const az = uniform(c, parent, kPhotonMuPair, kMuPairAzimuthDraw) * kPhotonTwoPi;
    double const k = (parent.energy_gev * 1000.) / (mu_mass * 1000.);
    double cosine = cos(1. / k);
    if (cosine == -1.) cosine *= -1.;
    Particle mu = child_of(parent, 13, secondary_energy(13, minus_total, mu_mass), parent.weight);
    Particle mp = child_of(parent, -13, secondary_energy(-13, plus_total, mu_mass), parent.weight);
    if (mu.energy_gev <= transport_mass(13)) mu.reserved = kNativeMuonPairCutFlag;
    if (mp.energy_gev <= transport_mass(-13)) mp.reserved = kNativeMuonPairCutFlag;
    deflect(parent.direction, cosine, az, mu.direction, false);
    deflect(parent.direction, cosine, fmod(az + kPhotonPi, kPhotonTwoPi), mp.direction, false);
    first_interaction(acc, parent, kPhotonMuPair, mu, &mp);
    mu.history_id = 0;
    emit(o, mu, kSlotOtherChild);
    mp.history_id = 1;
    emit(o, mp, kSlotOtherChild);
    ledger(acc, kMassCorrection, mass_correction(13, mu_mass, 2. * parent.weight));
    thinning_ledger(acc, 0, 2, 2, 0., 0.);
    return;
  }
  if (cap == kCapCompton) {
    double pc, ec;
    if (parent.generation == 0xffffffffU || parent.step_id == 0xffffffffffffffffULL ||
        !compton_cosines(parent.energy_gev, in.energy_fraction, pc, ec))
      return final_fallback(kFrInvalidFinalState);
    double const v = in.energy_fraction;
    double const az_u = uniform(c, parent, kCompton, 1);
    double const e1 = parent.energy_gev * (1. - v);
    double const e2 = secondary_energy(11, kElectronMassGeV + parent.energy_gev * v, kElectronMassGeV);
    Thin th;
    if (!two_child_thinning(c, parent, kCompton, e1, e2, false, th))
      return final_fallback(kFrInvalidFinalState);
    double const az = az_u * kPhotonTwoPi;
    Particle ph = child_of(parent, 22, e1, th.w1);
    deflect(parent.direction, pc, az, ph.direction, false);
    Particle el = child_of(parent, 11, e2, th.w2);
    deflect(parent.direction, ec, fmod(az + kPhotonPi, kPhotonTwoPi), el.direction, false);
    first_interaction(acc, parent, kCompton, ph, &el);
    std::uint32_t k = 0;
    if (th.keep & 1U) {
      ph.history_id = k++;
      emit(o, ph, kSlotNextChild);
    }
    if (th.keep & 2U) {
      el.history_id = k++;
      emit(o, el, kSlotOtherChild);
    }
    ledger(acc, kMassCorrection, mass_correction(11, kElectronMassGeV, (th.keep & 2U) ? th.w2 : 0.));
    ledger(acc, kMediumRestMass, kElectronMassGeV * w);
    thinning_ledger(acc, th.status, 2, k, (e1 + e2) * w,
                    ((th.keep & 1U) ? e1 * th.w1 : 0.) + ((th.keep & 2U) ? e2 * th.w2 : 0.));
    return;
  }
  if (cap == kCapPhotoelectric) {
    double etot, frac;
    if (parent.generation == 0xffffffffU || !isfinite(in.energy_fraction) ||
        fabs(in.energy_fraction - 1.) > 1.e-12 ||
        !photoelectric_energy(c.pair_lpm, in.component_hash, parent.energy_gev, etot, frac))
      return final_fallback(kFrInvalidFinalState);
    Particle el = child_of(parent, 11,
                           secondary_energy(11, kElectronMassGeV + parent.energy_gev * frac, kElectronMassGeV),
                           parent.weight);
    first_interaction(acc, parent, kPhotoelectric, el, nullptr);
    el.history_id = 0;
    emit(o, el, kSlotOtherChild);
    ledger(acc, kMassCorrection, mass_correction(11, kElectronMassGeV, parent.weight));
    ledger(acc, kMediumRestMass, kElectronMassGeV * w);
    ledger(acc, kPhotoelectricBinding, end.energy_gev * (1. - frac) * w);
    thinning_ledger(acc, 0, 1, 1, 0., 0.);
    return;
  }
  // photon pair
  if (!isfinite(parent.energy_gev) || parent.energy_gev < 2. * kElectronMassGeV ||
      parent.generation == 0xffffffffU || parent.step_id == 0xffffffffffffffffULL)
    return final_fallback(kFrInvalidFinalState);
  double const su = uniform(c, parent, kPhotonPair, 1);
  Tq const split = dq_loss_fraction(c, parent.pid, kPhotonPairFinalState, in.component_hash,
                                    parent.energy_gev * 1000., su);
  if (split.status != kTsSuccess) return final_fallback(fallback_reason(split.status));
  double rho = 0.;
  if (!decode_pair_split(parent.energy_gev, split.value, rho)) return final_fallback(kFrInvalidFinalState);
  double const ee = parent.energy_gev * rho;
  double const pe = parent.energy_gev - ee;
  if (!isfinite(rho) || !(rho > 0.) || !(rho < 1.) || ee < kElectronMassGeV || pe < kElectronMassGeV)
    return final_fallback(kFrInvalidFinalState);
  double const az_u = uniform(c, parent, kPhotonPair, 2);
  double const eu = uniform(c, parent, kPhotonPair, 3);
  double const pu = uniform(c, parent, kPhotonPair, 4);
  double survival = 0.;
  int const ls = pair_lpm(c.pair_lpm, in.component_hash, parent.energy_gev * 1000., rho, vertex_density, survival);
  if (ls != 0) return final_fallback(ls == 3 ? kFrInvalidDensity : kFrLpmUnavailable);
  double const lu = uniform(c, parent, kPhotonPair, 5);
  if (lu > survival) {
    Particle cont = parent;
    cont.step_id++;
    emit(o, cont, kSlotNextContinue);
    return;
  }
  double const e1 = secondary_energy(11, ee, kElectronMassGeV);
  double const e2 = secondary_energy(-11, pe, kElectronMassGeV);
  Thin th;
  if (!two_child_thinning(c, parent, kPhotonPair, e1, e2, false, th))
    return final_fallback(kFrInvalidFinalState);
  double const az = az_u * kPhotonTwoPi;
  Particle el = child_of(parent, 11, e1, th.w1);
  deflect(parent.direction, sauter_cosine(ee, eu), az, el.direction, false);
  Particle po = child_of(parent, -11, e2, th.w2);
  deflect(parent.direction, sauter_cosine(pe, pu), fmod(az + kPhotonPi, kPhotonTwoPi), po.direction, false);
  first_interaction(acc, parent, kPhotonPair, el, &po);
  std::uint32_t k = 0;
  if (th.keep & 1U) {
    el.history_id = k++;
    emit(o, el, kSlotOtherChild);
  }
  if (th.keep & 2U) {
    po.history_id = k++;
    emit(o, po, kSlotOtherChild);
  }
  double const kept = ((th.keep & 1U) ? th.w1 : 0.) + ((th.keep & 2U) ? th.w2 : 0.);
  ledger(acc, kMassCorrection, mass_correction(11, kElectronMassGeV, kept));
  thinning_ledger(acc, th.status, 2, k, (e1 + e2) * w,
                  ((th.keep & 1U) ? e1 * th.w1 : 0.) + ((th.keep & 2U) ? e2 * th.w2 : 0.));
}

// ---------------------------------------------------------------- lepton ----

// What the vertex part of an e+- step takes over from its transport part.
struct LeptonVertex {
  Particle parent;  // the particle at the vertex
  double total_rate;
  double process_uniform;
  double particle_mass_gev;
  double vertex_density;
};

// The move of an e+- step as the finish part needs it.
struct LeptonMove {
  Particle start;              // at the step start, layer resolved
  c8mi_particle_state end;     // before Molière scattering
  double continuous_deposit;
  double cut_deposit;
  double step_grammage;        // chord grammage of the step
  double vertex_density;
  double total_rate;
  double process_uniform;
  double particle_mass_gev;
  std::int32_t limit;
  std::int32_t process_id;
  std::uint32_t particle_cut;  // cut at the start: no scattering
  std::uint32_t observed_before_cut;
  std::uint32_t native_cut;    // a μ-pair muon cut at its first step (signed deposit, no track)
};

// The move of one e+- step (selection, continuous step, transport): its end
// state and limit. Returns false after recording a fallback.
C8MI_HD bool lepton_move(EmContext const& c, Accumulator const& acc, Particle const& p,
                         LeptonMove& mv) {
  Interaction in;
  Fallback fb{false, 0};
  select_interaction(c, p, in, fb);
  if (fb.flag) {
    record_fallback(acc, p, fb.reason, in.process_id, kStageSelection);
    return false;
  }
  auto transport_fallback = [&](int reason) {
    record_fallback(acc, in.particle, reason, in.process_id, kStageTransport);
    return false;
  };
  Particle start = in.particle;
  if (!is_charged_lepton(start.pid)) return transport_fallback(kFrUnsupportedParticle);
  if (in.status != kSelected && in.status != kDistanceSampled && in.status != kNoDiscrete &&
      in.status != kParticleCut)
    return transport_fallback(kFrInvalidFinalState);
  if ((in.status == kSelected || in.status == kDistanceSampled) &&
      (!isfinite(in.interaction_grammage) || in.interaction_grammage < 0.))
    return transport_fallback(kFrInvalidFinalState);
  if (in.status == kParticleCut && native_muon_pair_cut(start)) {
    // native_muon_pair_cut_record (NativeMuonPairCutRecord.inl:12-25): a
    // zero-length cut, no table query, no scattering; the deposit E - m is <= 0
    mv = LeptonMove{};
    mv.start = start;
    mv.end = start;
    mv.end.step_id++;
    mv.cut_deposit = start.energy_gev - transport_mass(start.pid);
    mv.limit = kLlCut;
    mv.particle_cut = 1u;
    mv.native_cut = 1u;
    return true;
  }
  // prepare_lepton_continuous_step (LeptonContinuousStep.inl:43-120)
  Tq const native_mass = cq_mass(c, start.pid);
  double const tm = 1000. * transport_mass(start.pid);
  if (native_mass.status != kTsSuccess) return transport_fallback(fallback_reason(native_mass.status));
  if (!(tm > 0.)) return transport_fallback(fallback_reason(kTsContParticle));
  Tq const minimum = cq_minimum_energy(c, start.pid);
  if (minimum.status != kTsSuccess) return transport_fallback(fallback_reason(minimum.status));
  double const e0 = start.energy_gev * 1000.;
  if (!(e0 > tm)) return transport_fallback(kFrInvalidFinalState);
  Tq const tcut = cq_transport_cut(c, start.pid);
  if (tcut.status != kTsSuccess) return transport_fallback(kFrInvalidQuery);
  // muons with continuous tables: this backend has no muon decay or muon
  // scattering, so they stay with the host
  if (!is_ee(start.pid)) return transport_fallback(kFrUnsupportedParticle);
  bool const particle_cut = in.status == kParticleCut || e0 - tm < tcut.value;
  double max_grammage = 0.;
  if (!particle_cut) {
    Tq const r0 = native(range(c.tables, start.pid, e0), true);
    if (r0.status != kTsSuccess) return transport_fallback(fallback_reason(r0.status));
    double const rl = (1. - kMaxRelativeLoss) * e0;
    double const target = rl > minimum.value ? rl : minimum.value;
    Tq const r1 = native(range(c.tables, start.pid, target), true);
    if (r1.status != kTsSuccess) return transport_fallback(fallback_reason(r1.status));
    max_grammage = r0.value - r1.value;
    double const scale = r0.value > 1. ? r0.value : 1.;
    if (max_grammage < 0. && max_grammage > -64. * 2.22044604925031308085e-16 * scale) max_grammage = 0.;
    if (!isfinite(max_grammage) || max_grammage < 0.) return transport_fallback(kFrInvalidQuery);
  }
  env::Layer const layer = env::query(c, start.position_m, start.direction, start.environment_node);
  if (layer.status != env::kSuccess) return transport_fallback(kFrUnsupportedGeometry);
  start.environment_node = layer.node;
  Particle end = start;
  int limit;
  double continuous_deposit = 0., cut_deposit = 0.;
  bool observed_before_cut = false;
  double vertex_density = 0.;
  double step_grammage = 0.;
  if (particle_cut) {
    end.step_id++;
    limit = kLlCut;
    cut_deposit = (e0 - tm) / 1000.;
  } else {
    // maximum_uniform_magnetic_step, lepton_boundary, lepton_observation_distance
    // (LeptonTransportStep.inl:119-160)
    double const charge = start.pid > 0 ? -1. : 1.;
    double const* field = env::model_of(c, layer.node)->magnetic_field_t;
    double const m_gev = tm / 1000.;
    mag::Limit const mlimit = mag::step_limit(start, m_gev, charge, field, c.maximum_magnetic_deflection);
    if (mlimit.status == mag::kInvalidInput || mlimit.status == mag::kNonFinite)
      return transport_fallback(kFrMagneticTransport);
    double geometry;
    std::int32_t next_node;
    if (!mag::lepton_boundary(c, start, m_gev, charge, field, mlimit, geometry, next_node))
      return transport_fallback(kFrMagneticBoundary);
    bool magnetic_wins = next_node == start.environment_node;
    bool obs_wins = false;
    if (c.observation_enabled) {
      double const od = mag::lepton_observation_distance(c, start, m_gev, charge, field, mlimit);
      if (od < geometry) {
        geometry = od;
        obs_wins = true;
      }
      if (obs_wins) magnetic_wins = false;
    }
    env::Grammage const gg = env::competition_grammage(c, layer.node, start.position_m, start.direction, geometry);
    if (gg.status != env::kSuccess) return transport_fallback(kFrGrammage);
    bool const has_interaction = in.status == kSelected || in.status == kDistanceSampled;
    double const ig = has_interaction ? in.interaction_grammage : __builtin_huge_val();
    bool const continuous_wins = max_grammage < gg.value && max_grammage < ig;
    bool const interaction_wins = !continuous_wins && has_interaction && ig <= gg.value;
    double traversed = gg.value;
    double distance = geometry;
    if (continuous_wins) traversed = max_grammage;
    else if (interaction_wins) traversed = ig;
    if (continuous_wins || interaction_wins) {
      env::Grammage const d = env::arclength(c, layer.node, start.position_m, start.direction, traversed);
      if (d.status != env::kSuccess || d.value > geometry * (1. + 1.e-12))
        return transport_fallback(kFrInverseGrammage);
      distance = d.value;
    }
    bool const observation_reached = obs_wins && !continuous_wins && !interaction_wins && !magnetic_wins;
    // advance_uniform_magnetic_field, with no field where the limit was Linear
    double const zero_field[3] = {0., 0., 0.};
    int const advanced = mag::advance(start, m_gev, charge, mlimit.status == mag::kLinear ? zero_field : field,
                                      distance, end);
    if (advanced == mag::kInvalidInput || advanced == mag::kNonFinite) return transport_fallback(kFrMagneticTransport);
    double chord_dir[3], chord2 = 0.;
    for (int a = 0; a < 3; ++a) {
      double const dd = end.position_m[a] - start.position_m[a];
      chord_dir[a] = dd;
      chord2 += dd * dd;
    }
    double const chord = sqrt(chord2);
    if (chord > 0.) {
      for (int a = 0; a < 3; ++a) chord_dir[a] /= chord;
    } else {
      for (int a = 0; a < 3; ++a) chord_dir[a] = start.direction[a];
    }
    env::Grammage const cg = env::grammage(c, layer.node, start.position_m, chord_dir, chord);
    if (cg.status != env::kSuccess) return transport_fallback(kFrGrammage);
    traversed = cg.value;
    step_grammage = traversed;
    Tq const fe = cq_energy_after_loss(c, start.pid, e0, traversed);
    double final_mev;
    bool cut_reached = false;
    if (fe.status == kTsSuccess) {
      final_mev = fe.value;
    } else if (fe.status == kTsTransportCut) {
      final_mev = minimum.value;
      cut_reached = true;
    } else {
      return transport_fallback(fallback_reason(fe.status));
    }
    double const loss = e0 - final_mev;
    if (continuous_wins && !(loss > 0.)) return transport_fallback(kFrInvalidQuery);
    end.energy_gev = final_mev / 1000.;
    continuous_deposit = loss / 1000.;
    if (end.time_s > kCutTime) {
      end.step_id++;
      limit = kLlCut;
      cut_deposit = (final_mev - tm) / 1000.;
      observed_before_cut = observation_reached;
    } else if (cut_reached || final_mev - tm < tcut.value) {
      end.step_id++;
      limit = kLlCut;
      if (!observation_reached) {
        env::Layer const el = env::query(c, end.position_m, end.direction, start.environment_node);
        if (el.status != env::kSuccess) return transport_fallback(kFrUnsupportedGeometry);
      }
      cut_deposit = (final_mev - tm) / 1000.;
      observed_before_cut = observation_reached;
    } else if (interaction_wins) {
      env::Layer const v = env::query(c, end.position_m, end.direction, start.environment_node);
      if (v.status != env::kSuccess || !isfinite(v.density) || !(v.density > 0.))
        return transport_fallback(kFrVertexLookup);
      limit = kLlInteraction;
      vertex_density = v.density;
    } else {
      end.step_id++;
      if (continuous_wins) {
        env::Layer const el = env::query(c, end.position_m, end.direction, start.environment_node);
        if (el.status != env::kSuccess) return transport_fallback(kFrUnsupportedGeometry);
        limit = kLlContinuous;
      } else if (magnetic_wins) {
        env::Layer const el = env::query(c, end.position_m, end.direction, start.environment_node);
        if (el.status != env::kSuccess) return transport_fallback(kFrMagneticBoundary);
        limit = kLlMagnetic;
      } else if (observation_reached) {
        limit = kLlObservation;
      } else {
        end.environment_node = next_node;
        c8mi_env_model const* nm = env::model_of(c, next_node);
        if (!nm) {
          limit = kLlEscaped;
        } else {
          env::Layer const next = env::query(c, end.position_m, end.direction, next_node);
          if (next.status != env::kSuccess || next_node == start.environment_node)
            return transport_fallback(kFrUnsupportedGeometry);
          limit = kLlLayer;
          end.medium_id = nm->medium_id;
        }
      }
    }
  }
  mv.start = start;
  mv.end = end;
  mv.continuous_deposit = continuous_deposit;
  mv.cut_deposit = cut_deposit;
  mv.step_grammage = step_grammage;
  mv.vertex_density = vertex_density;
  mv.total_rate = in.total_rate;
  mv.process_uniform = in.process_uniform;
  mv.particle_mass_gev = in.particle_mass_gev;
  mv.limit = limit;
  mv.process_id = in.process_id;
  mv.particle_cut = particle_cut ? 1u : 0u;
  mv.observed_before_cut = observed_before_cut ? 1u : 0u;
  mv.native_cut = 0u;
  return true;
}

// apply_moliere_scattering_stage (MoliereStep.inl:74-99) on a move, with
// sample_moliere_scattering_angle_2d (MoliereScattering.inl:513-540): draws
// keyed by the start state, two projected angles, and on success the end
// direction becomes normalize(tracked + scattered - start). A failed sampling
// or deflection turns the step into a transport fallback. `Table`: the
// interpolation view's sampler, else the analytical functions.
enum MoliereOutcome : int { kMoliereDone = 0, kMoliereFailed = 1, kMoliereOpen = 2 };

// A sample whose Newton iteration did not converge in the step kernel: where
// to continue (stage 1: the first projected angle, 2: the second).
struct MoliereOpen {
  std::uint32_t source;
  std::uint32_t stage;
  std::int32_t iterations;  // done so far for this sample
  std::uint32_t pad;        // the tail kernel's outcome (a front source state) for the open finish
  double current;           // the Newton iterate
  double first_angle;       // stage 2: the first projected angle
};

struct MoliereDraws {
  double u1, u2, ua;
};

C8MI_HD MoliereDraws moliere_draws(EmContext const& c, Particle const& start) {
  return {uniform(c, start, kScatteringProcess, 0), uniform(c, start, kScatteringProcess, 1),
          uniform(c, start, kScatteringProcess, 2)};
}

// The space angle of the two projected angles, then the deflection of the end direction.
C8MI_HD int moliere_apply(LeptonMove& mv, double a1, double a2, double ua) {
  double const angle = sqrt(a1 * a1 + a2 * a2);
  double dir[3];
  if (!isfinite(angle) || moliere::deflect_end(mv.start.direction, mv.end.direction, angle, ua, dir) != 0)
    return kMoliereFailed;
  for (int a = 0; a < 3; ++a) mv.end.direction[a] = dir[a];
  return kMoliereDone;
}

// Up to the context's moliere_fast_iterations Newton iterations per sample; a
// sample still open is described in `open` (kMoliereOpen) for
// moliere_scatter_wave.
template <bool Table>
C8MI_HD int moliere_scatter(EmContext const& c, LeptonMove& mv, MoliereOpen& open) {
  if (mv.particle_cut) return kMoliereDone;
  MoliereDraws const d = moliere_draws(c, mv.start);
  moliere::Quotients const q{c.moliere_i1, c.moliere_i2, c.moliere_i2_large, c.moliere_half_over};
  moliere::Table const t{c.moliere_polynomials, c.moliere_guess};
  moliere::State<moliere::kDeviceComponents> st;
  int status = moliere::build(c.moliere, mv.step_grammage, mv.start.energy_gev * 1000., st);
  if (status == moliere::kNoDeflection) return kMoliereDone;
  if (status != moliere::kSuccess) return kMoliereFailed;
  double a[2];
  double const u[2] = {d.u1, d.u2};
  for (int k = 0; k < 2; ++k) {
    status = moliere::newton_start<Table>(c.moliere, t, st, u[k], a[k]);
    int const fast = static_cast<int>(c.moliere_fast_iterations);
    if (status == moliere::kStarted)
      status = moliere::newton_iterate<Table>(c.moliere, q, t, st, u[k], 1, fast, a[k]);
    if (status == moliere::kPending) {
      open = {0u, static_cast<std::uint32_t>(k + 1), fast, 0u, a[k], k ? a[0] : 0.};
      return kMoliereOpen;
    }
    if (status != moliere::kSuccess) return kMoliereFailed;
  }
  return moliere_apply(mv, a[0], a[1], d.ua);
}

// The rest of an open sample (and the second one after it) on one workgroup of
// moliere::kTailThreads threads; every thread returns the same outcome and
// deflected move.
template <bool Table>
__device__ __forceinline__ int moliere_scatter_wave(EmContext const& c, LeptonMove& mv, MoliereOpen const& open,
                                                    double* jobs) {
  MoliereDraws const d = moliere_draws(c, mv.start);
  moliere::Quotients const q{c.moliere_i1, c.moliere_i2, c.moliere_i2_large, c.moliere_half_over};
  moliere::Table const t{c.moliere_polynomials, c.moliere_guess};
  moliere::State<moliere::kDeviceComponents> st;
  moliere::build(c.moliere, mv.step_grammage, mv.start.energy_gev * 1000., st);  // succeeded before
  double a1, a2;
  int status;
  if (open.stage == 1) {
    a1 = open.current;
    status = moliere::newton_iterate_wave<Table>(c.moliere, q, t, st, d.u1, open.iterations + 1, jobs, a1);
    if (status != moliere::kSuccess) return kMoliereFailed;
    status = moliere::newton_start<Table>(c.moliere, t, st, d.u2, a2);
    if (status == moliere::kStarted)
      status = moliere::newton_iterate_wave<Table>(c.moliere, q, t, st, d.u2, 1, jobs, a2);
  } else {
    a1 = open.first_angle;
    a2 = open.current;
    status = moliere::newton_iterate_wave<Table>(c.moliere, q, t, st, d.u2, open.iterations + 1, jobs, a2);
  }
  if (status != moliere::kSuccess) return kMoliereFailed;
  return moliere_apply(mv, a1, a2, d.ua);
}

// The rest of the step after Molière scattering (applied to `mv` beforehand, or
// off): profile and ledger, records and outputs. `moliere_failed`: the
// scattering failed, the step becomes a transport fallback whose state is the
// transport record's interaction state (the vertex for an interaction
// candidate, else the step's input particle, `input()`, fetched only then).
// Returns true when the step ends at an interaction vertex: `vx` then holds
// what lepton_vertex needs.
template <class Input>
C8MI_HD bool lepton_finish(EmContext const& c, Accumulator const& acc, Input const& input,
                           LeptonMove const& mv, bool moliere_failed, Outputs& o,
                           LeptonVertex& vx) {
  Particle const& start = mv.start;
  Particle end;
  static_cast<c8mi_particle_state&>(end) = mv.end;
  end.shower_key = start.shower_key;
  int const limit = mv.limit;
  double const continuous_deposit = mv.continuous_deposit, cut_deposit = mv.cut_deposit;
  bool const observed_before_cut = mv.observed_before_cut != 0;

  if (moliere_failed) {
    if (limit == kLlInteraction) record_fallback(acc, end, kFrMoliereFailed, 0, kStageTransport);
    else record_fallback(acc, input(), kFrMoliereFailed, mv.process_id, kStageTransport);
    return false;
  }
  o.step = true;
  double const w = start.weight;
  double const tmg = transport_mass(start.pid);
  if (mv.native_cut) {
    // accumulate_lepton_profile_step of the native cut (ProfileAccumulation.inl:300-345):
    // no track, a zero-length profile step, the signed deposit into the muon
    // energy loss (its own ledger term here), the rest mass as at any cut
    if (o.track) write_no_track(o.track, o.track_stride);
    double const deposit = cut_deposit * w;
    if (deposit < 0. && c.profile_bins != 0) count(acc, kDepositedSteps);
    ledger(acc, kNativeMuonCut, deposit);
    count(acc, kSteps);
    count(acc, kLeptonLimit0 + kLlCut);
    ledger(acc, kCutRestMass, tmg * w);
    return false;
  }
  if (o.track) {
    write_track(o.track, o.track_stride, start, end);  // e+- only reach here
    o.tracked = true;
  }
  // accumulate_lepton_profile_step
  double const gs = project_grammage(acc, start.position_m);
  double const ge = project_grammage(acc, end.position_m);
  particle_profile(acc, start.pid, gs, ge, w);
  energy_profile(acc, gs, ge, continuous_deposit * w);
  point_energy_profile(acc, ge, cut_deposit * w);
  ledger(acc, kContinuous, continuous_deposit * w);
  count(acc, kSteps);
  count(acc, kLeptonLimit0 + limit);
  if (limit == kLlCut) {
    ledger(acc, kCutDeposit, cut_deposit * w);
    ledger(acc, kCutRestMass, tmg * w);
    if (observed_before_cut) {
      ledger(acc, kOverlap, (cut_deposit + tmg) * w);
      terminal_energy(acc, end.energy_gev * w, true);
      record_observation(acc, end);
    }
    return false;
  }
  if (limit == kLlObservation || limit == kLlEscaped) {
    terminal_energy(acc, end.energy_gev * w, limit == kLlObservation);
    record_observation(acc, end);
    return false;
  }
  if (limit != kLlInteraction) {
    emit(o, end, kSlotNextContinue);
    return false;
  }
  o.interaction = true;
  vx.parent = end;
  vx.total_rate = mv.total_rate;
  vx.process_uniform = mv.process_uniform;
  vx.particle_mass_gev = mv.particle_mass_gev;
  vx.vertex_density = mv.vertex_density;
  return true;
}

// The step of one e+- up to its interaction vertex (move, then finish),
// without Molière scattering.
C8MI_HD bool lepton_transport(EmContext const& c, Accumulator const& acc, Particle const& p,
                              Outputs& o, LeptonVertex& vx) {
  LeptonMove mv;
  if (!lepton_move(c, acc, p, mv)) return false;
  return lepton_finish(c, acc, [&] { return p; }, mv, false, o, vx);
}

// Interaction vertex of an e+- step: vertex reselection at the post-loss
// energy, then the final state.
C8MI_HD void lepton_vertex(EmContext const& c, Accumulator const& acc, LeptonVertex const& vx,
                           Outputs& o) {
  // select_lepton_vertex (LeptonVertexSelection.inl:41-197)
  Particle parent = vx.parent;
  double const w = parent.weight;
  std::int32_t process = 0;
  auto vertex_fallback = [&](int reason) {
    record_fallback(acc, parent, reason, process, kStageVertex);
  };
  if (!isfinite(vx.total_rate) || !(vx.total_rate > 0.) || !isfinite(vx.process_uniform) ||
      !(vx.process_uniform > 0.) || !(vx.process_uniform < 1.))
    return vertex_fallback(kFrInvalidFinalState);
  double const e_mev = parent.energy_gev * 1000.;
  double const outer = vx.process_uniform * vx.total_rate;
  Tq const vt = native(total_rate(c.tables, parent.pid, e_mev), false);
  if (vt.status != kTsSuccess) return vertex_fallback(fallback_reason(vt.status));
  if (!(vt.value > 0.) || outer >= vt.value) {
    if (parent.step_id == 0xffffffffffffffffULL) return vertex_fallback(kFrInvalidFinalState);
    Particle cont = parent;
    cont.step_id++;
    emit(o, cont, kSlotNextContinue);
    return;
  }
  double const us = uniform(c, parent, kSelectionProcess, 0);
  Selection const sel = select_column(c, parent.pid, e_mev, us);
  if (sel.status != kTsSuccess) {
    process = sel.process;
    return vertex_fallback(fallback_reason(sel.status));
  }
  if (!sel.selected) return vertex_fallback(kFrZeroRate);
  process = sel.process;
  if (requires_replay(parent.pid, sel.process, sel.quantile)) return vertex_fallback(kFrSelectionReplay);
  Tq const loss = dq_loss_fraction(c, parent.pid, sel.process, sel.component, e_mev, sel.quantile);
  if (loss.status != kTsSuccess) return vertex_fallback(fallback_reason(loss.status));
  double const v = loss.value;
  // classify_lepton_final_state / materialize (LeptonFinalStateStep.inl:322-1036)
  auto final_fallback = [&](int reason) {
    record_fallback(acc, parent, reason, process, kStageFinal);
  };
  int const cap = capability(parent.pid, process);
  if (cap != kCapBrems && cap != kCapAnnihilation && cap != kCapIonization)
    return final_fallback(cap == kCapCpuOnly ? kFrCpuOnly : kFrNotImplemented);
  c8mi_brems_lpm const& lpm = c.brems_lpm;
  double const me = lpm.electron_mass_mev / 1000.;
  double const lm = vx.particle_mass_gev > 0. ? vx.particle_mass_gev : lpm.lepton_mass_mev / 1000.;
  double const fs_me = lpm.lepton_mass_mev / 1000.;
  if (!isfinite(parent.energy_gev) || !(parent.energy_gev > lm) || parent.generation == 0xffffffffU ||
      parent.step_id == 0xffffffffffffffffULL)
    return final_fallback(kFrInvalidFinalState);
  if (cap == kCapAnnihilation) {
    if (!isfinite(v) || fabs(v - 1.) > 1.e-12) return final_fallback(kFrInvalidFinalState);
    double const ru = uniform(c, parent, kAnnihilation, 1);
    double const au = uniform(c, parent, kAnnihilation, 2);
    double rho;
    if (!annihilation_rho(parent.energy_gev, lm, ru, rho)) return final_fallback(kFrInvalidFinalState);
    double const total = parent.energy_gev + lm;
    Thin th;
    if (!two_child_thinning(c, parent, kAnnihilation, total * (1. - rho), total * rho, true, th))
      return final_fallback(kFrInvalidFinalState);
    double const e1 = total * (1. - rho), e2 = total * rho;
    double const mom = sqrt((parent.energy_gev + lm) * (parent.energy_gev - lm));
    double const c1 = (total * (1. - rho) - lm) / ((1. - rho) * mom);
    double const c2 = (total * rho - lm) / (rho * mom);
    if (!isfinite(e1) || !isfinite(e2) || !(e1 > 0.) || !(e2 > 0.) || !isfinite(c1) || !isfinite(c2)) {
      count(acc, kInvalid);
      return;
    }
    double const az = au * kLeptonTwoPi;
    Particle g1 = child_of(parent, 22, e1, th.w1);
    Particle g2 = child_of(parent, 22, e2, th.w2);
    deflect(parent.direction, c1, az, g1.direction, true);
    deflect(parent.direction, c2, fmod(az + kLeptonPi, kLeptonTwoPi), g2.direction, true);
    first_interaction(acc, parent, kAnnihilation, g1, &g2);
    std::uint32_t k = 0;
    if (th.keep & 1U) {
      g1.history_id = k++;
      emit(o, g1, kSlotOtherChild);
    }
    if (th.keep & 2U) {
      g2.history_id = k++;
      emit(o, g2, kSlotOtherChild);
    }
    ledger(acc, kMediumRestMass, kElectronMassGeV * w);
    thinning_ledger(acc, th.status, 2, k, (e1 + e2) * w,
                    ((th.keep & 1U) ? e1 * th.w1 : 0.) + ((th.keep & 2U) ? e2 * th.w2 : 0.));
    return;
  }
  if (cap == kCapIonization) {
    double const ev = (v * (parent.energy_gev * 1000.)) / (parent.energy_gev * 1000.);
    double const outgoing = parent.energy_gev * (1. - ev);
    double const delta = parent.energy_gev * ev + me;
    if (!isfinite(ev) || !(ev > 0.) || !(ev < 1.) || !isfinite(outgoing) || outgoing < lm ||
        !isfinite(delta) || !(delta > me))
      return final_fallback(kFrInvalidFinalState);
    double const au = uniform(c, parent, kIonization, 1);
    double const e1 = secondary_energy(parent.pid, outgoing, lm);
    double const e2 = secondary_energy(11, delta, me);
    Thin th;
    if (!two_child_thinning(c, parent, kIonization, e1, e2, true, th)) return final_fallback(kFrInvalidFinalState);
    double const oe = parent.energy_gev * (1. - ev);
    double const de = parent.energy_gev * ev + fs_me;
    double const im = sqrt((parent.energy_gev + lm) * (parent.energy_gev - lm));
    double const om = sqrt((oe + lm) * (oe - lm));
    double const dm = sqrt((de + fs_me) * (de - fs_me));
    double const oc = ((parent.energy_gev + fs_me) * oe - parent.energy_gev * fs_me - lm * lm) / (im * om);
    double const dc = ((parent.energy_gev + fs_me) * de - parent.energy_gev * fs_me - fs_me * fs_me) / (im * dm);
    if (!isfinite(oc) || !isfinite(dc)) {
      count(acc, kInvalid);
      return;
    }
    double const az = au * kLeptonTwoPi;
    Particle out = child_of(parent, parent.pid, secondary_energy(parent.pid, oe, lm), th.w1);
    Particle dl = child_of(parent, 11, secondary_energy(11, de, fs_me), th.w2);
    deflect(parent.direction, oc, az, out.direction, true);
    deflect(parent.direction, dc, fmod(az + kLeptonPi, kLeptonTwoPi), dl.direction, true);
    first_interaction(acc, parent, kIonization, out, &dl);
    std::uint32_t k = 0;
    if (th.keep & 1U) {
      out.history_id = k++;
      emit(o, out, kSlotNextChild);
    }
    if (th.keep & 2U) {
      dl.history_id = k++;
      emit(o, dl, kSlotNextChild);
    }
    ledger(acc, kMassCorrection,
           mass_correction(parent.pid, lm, (th.keep & 1U) ? th.w1 : 0.) +
               mass_correction(11, fs_me, (th.keep & 2U) ? th.w2 : 0.));
    ledger(acc, kMediumRestMass, kElectronMassGeV * w);
    thinning_ledger(acc, th.status, 2, k, (out.energy_gev + dl.energy_gev) * w,
                    ((th.keep & 1U) ? out.energy_gev * th.w1 : 0.) +
                        ((th.keep & 2U) ? dl.energy_gev * th.w2 : 0.));
    return;
  }
  // bremsstrahlung
  double const pe = parent.energy_gev * v;
  double const le = parent.energy_gev - pe;
  if (!isfinite(v) || !(v > 0.) || !(v < 1.) || !isfinite(pe) || !(pe > 0.) || !isfinite(le) || le < lm)
    return final_fallback(kFrInvalidFinalState);
  double const au = uniform(c, parent, kBrems, 1);
  double survival = 0.;
  // the vertex density comes from the transport step
  int const ls = brems_lpm(lpm, sel.component, parent.energy_gev * 1000., v, vx.vertex_density, survival);
  if (ls != 0) return final_fallback(ls == 3 ? kFrInvalidDensity : kFrLpmUnavailable);
  double const lu = uniform(c, parent, kBrems, 2);
  if (lu > survival) {
    Particle cont = parent;
    cont.step_id++;
    emit(o, cont, kSlotNextContinue);
    return;
  }
  double const e1 = secondary_energy(parent.pid, le, lm);
  Thin th;
  if (!two_child_thinning(c, parent, kBrems, e1, pe, true, th)) return final_fallback(kFrInvalidFinalState);
  double const az = au * kLeptonTwoPi;
  double pdir[3];
  deflect(parent.direction, cos(lm / parent.energy_gev), az, pdir, true);
  Particle lep = child_of(parent, parent.pid, e1, th.w1);
  double const pm2 = (parent.energy_gev + lm) * (parent.energy_gev - lm);
  double const pmom = sqrt(0. > pm2 ? 0. : pm2);
  for (int a = 0; a < 3; ++a) lep.direction[a] = parent.direction[a] * pmom - pdir[a] * pe;
  double const n2 = env::dot3(lep.direction, lep.direction);
  if (!isfinite(n2) || !(n2 > 0.)) {
    count(acc, kInvalid);
    return;
  }
  double const inv = 1. / sqrt(n2);
  for (int a = 0; a < 3; ++a) lep.direction[a] *= inv;
  if (!isfinite(lep.direction[0]) || !isfinite(lep.direction[1]) || !isfinite(lep.direction[2])) {
    count(acc, kInvalid);
    return;
  }
  Particle ph = child_of(parent, 22, pe, th.w2);
  for (int a = 0; a < 3; ++a) ph.direction[a] = pdir[a];
  first_interaction(acc, parent, kBrems, lep, &ph);
  std::uint32_t k = 0;
  if (th.keep & 1U) {
    lep.history_id = k++;
    emit(o, lep, kSlotNextChild);
  }
  if (th.keep & 2U) {
    ph.history_id = k++;
    emit(o, ph, kSlotOtherChild);
  }
  ledger(acc, kMassCorrection, mass_correction(parent.pid, lm, (th.keep & 1U) ? th.w1 : 0.));
  thinning_ledger(acc, th.status, 2, k, (e1 + pe) * w,
                  ((th.keep & 1U) ? e1 * th.w1 : 0.) + ((th.keep & 2U) ? pe * th.w2 : 0.));
}

}  // namespace c8mi::em::phys
