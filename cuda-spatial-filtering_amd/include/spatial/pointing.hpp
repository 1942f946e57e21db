// Array geometry and beam-target types that the pipelines' BeamSteering
// constructor takes (include/spatial/pointing.hpp:27-61 in the reference).
// They carry the same fields so that callers build steering objects
// unchanged; the pointing computations themselves (RA/Dec -> direction
// cosines -> steering phases) are not part of this path.
#pragma once
#include <string>

// Antenna position relative to the array reference point, local East-North-Up
// frame, metres.
struct ENUPosition {
  double east = 0.0;
  double north = 0.0;
  double up = 0.0;
};

// Array reference point (WGS84).
struct ArrayLocation {
  double latitude_deg = 0.0;
  double longitude_deg = 0.0;
  double height_m = 0.0;
};

// Linear channel-to-frequency map: sky frequency of absolute channel 0 and the
// channel width.
struct FrequencyPlan {
  double base_frequency_hz = 0.0;
  double channel_bandwidth_hz = 0.0;
};

// A beam's pointing target: "radec" (J2000) or "zenith".
struct BeamTarget {
  std::string mode = "zenith";
  double ra_deg = 0.0;
  double dec_deg = 0.0;
};

inline double channel_to_frequency_hz(int absolute_channel_index, const FrequencyPlan &plan) {
  return plan.base_frequency_hz + static_cast<double>(absolute_channel_index) * plan.channel_bandwidth_hz;
}
