"""rbps_amd: MI355X-native XPBD substep pipeline behind the rbps World API."""
