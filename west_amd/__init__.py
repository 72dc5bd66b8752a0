"""west_amd: MI355X-native Wstat dielectric-response hot path."""
