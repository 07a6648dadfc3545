// Test mock (the reference's project-wide switches are not needed here).
#pragma once
