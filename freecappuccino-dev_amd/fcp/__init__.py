"""fcp: MI355X-native finite-volume assemble-and-solve path (host side)."""
